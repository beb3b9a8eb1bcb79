"""Writes the golden fixtures of the reference's example notebook (examples/example_diffractom_pipeline.ipynb):

  example_data_nz.npz   the measured data of examples/example_data.h5 (180 x 103 x 180 x 5 float32) as its non-zero
                        entries (delta-coded flat indices + values, lossless), with the file's lattice, hkl list and
                        wavelength
  apriori_grid.npy      examples/apriori_grid.npy (543 orientation matrices), copied
  golden_objectives.json  the SPDHG objectives the notebook printed (cells 13 and 15) and their set-up

The HDF5 file is read without h5py: every dataset of it is stored contiguously, and its data-layout messages
(version 3, class 1: address and size, little-endian uint64) give where; datasets are told apart by their sizes.
Run in the container that holds the reference: python tests/golden/make_fixtures.py
"""
import json
import re
import struct
from pathlib import Path

import numpy as np

REF = Path("/root/reference/examples")
OUT = Path(__file__).resolve().parent
SHAPE = (180, 103, 180, 5)


def contiguous_datasets(raw):
    """{size: address} of the contiguous-layout messages of an HDF5 file."""
    out = {}
    for m in re.finditer(re.escape(b"\x03\x01"), raw):
        i = m.end()
        if i + 16 > len(raw):
            continue
        addr, size = struct.unpack_from("<QQ", raw, i)
        if 0 < size < len(raw) and 0 < addr < len(raw) and addr + size <= len(raw):
            out.setdefault(size, addr)
    return out


def main():
    raw = (REF / "example_data.h5").read_bytes()
    ds = contiguous_datasets(raw)
    n = int(np.prod(SHAPE))
    data = np.frombuffer(raw, dtype="<f4", count=n, offset=ds[4 * n]).reshape(SHAPE)
    lattice = np.frombuffer(raw, dtype="<f8", count=9, offset=ds[72]).reshape(3, 3)
    hkl = np.frombuffer(raw, dtype="<i8", count=15, offset=ds[120]).reshape(5, 3)
    assert np.isfinite(data).all() and data.min() >= 0
    assert np.allclose(lattice, np.diag([lattice[0, 0]] * 3)) and hkl.tolist()[0] == [1, 1, 1]
    # the wavelength: the 8-byte dataset whose value the notebook printed (0.24797 A, cell 3)
    lam = [struct.unpack_from("<d", raw, a)[0] for s, a in ds.items() if s == 8]
    lam = [v for v in lam if 0.2 < v < 0.3]
    assert len(lam) == 1
    flat = data.ravel()
    nz = np.nonzero(flat)[0].astype(np.uint32)
    didx = np.diff(nz, prepend=np.uint32(0)).astype(np.uint32)
    np.savez_compressed(OUT / "example_data_nz.npz", didx=didx, val=flat[nz], shape=np.array(SHAPE),
                        lattice=lattice, hkl=hkl, wavelength=np.array(lam[0]))
    np.save(OUT / "apriori_grid.npy", np.load(REF / "apriori_grid.npy", allow_pickle=False))
    golden = {
        "source": "examples/example_diffractom_pipeline.ipynb cells 5, 9, 11, 13, 15 (printed outputs)",
        "cfg": {"Nx": 99, "Ny": 99, "My": 103, "N_Omega": 180, "angle_range": [0, 180], "cor_offset": 0,
                "N_eta": 180, "eta_angle_range": [0, 360], "k_direction_0": [0, 0, 1], "p_direction_0": [1, 0, 0],
                "detector_direction_origin": [0, -1, 0], "detector_direction_positive_90": [0, 0, -1]},
        "energy": "12.398 / wavelength",
        "weights_zero_eta_bins": [[40, 50], [130, 140]],
        "solver": {"subsets": 8, "seed": 0, "n_epochs": 30, "diagnostics_interval": 10},
        "tailored": {"grid": "apriori_grid.npy", "sigma_deg": 0.5, "K": 543,
                     "f": {"10": 1.14836889e6, "20": 1.13798917e6, "30": 1.10145372e6}},
        "uniform": {"grid": "np.random.seed(0); Grid.from_random_fundamental_zone(50000, 'cubic', 2 deg); "
                            "prune_close_orientations(theta_deg=3.0, target=15000)", "sigma_deg": 2.0, "K": 9899,
                    "f": {"10": 7.05589124e6, "20": 5.74759022e6, "30": 5.54666584e6}},
        "reflections": {"hkl": [[-1, -1, -1], [-2, 0, 0], [-2, -2, 0], [-3, -1, -1], [-2, -2, -2]],
                        "multiplicity": [8, 6, 12, 24, 8],
                        "d_spacing": [2.33786445, 2.02465, 1.43164374, 1.22090989, 1.16893222],
                        "two_theta": [0.10611598, 0.12255137, 0.17342262, 0.20345202, 0.21253218],
                        "source": "notebook cell 7 output"},
    }
    (OUT / "golden_objectives.json").write_text(json.dumps(golden, indent=1))
    print("wrote", sorted(p.name for p in OUT.iterdir()))


if __name__ == "__main__":
    main()
