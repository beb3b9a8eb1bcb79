"""The HDF5 reader (diffractom_amd.io) on the reference's example data file, where that file is present (this
container; the GPU box has no reference tree): every dataset the notebook reads equals the committed fixture
(tests/golden/example_data_nz.npz, written by tests/golden/make_fixtures.py from the file's raw layout blocks)."""
from pathlib import Path

import numpy as np
import pytest

H5 = Path("/root/reference/examples/example_data.h5")
FIX = Path(__file__).resolve().parent / "golden" / "example_data_nz.npz"


@pytest.mark.skipif(not H5.exists(), reason="the reference's example data file is not present")
def test_example_data_matches_the_fixture():
    from diffractom_amd.io import load_example_data, read_hdf5
    d = read_hdf5(H5)
    assert {"data_array", "lattice", "hkl_list", "wavelength"} <= set(d)
    data, lattice, hkl, lam = load_example_data(H5)
    z = np.load(FIX)
    flat = np.zeros(int(np.prod(z["shape"])), np.float32)
    flat[np.cumsum(z["didx"].astype(np.int64))] = z["val"]
    assert data.shape == tuple(z["shape"]) and data.dtype == np.float32
    assert np.array_equal(data.ravel(), flat)
    assert np.array_equal(lattice, z["lattice"]) and np.array_equal(hkl, z["hkl"])
    assert lam == float(z["wavelength"])
    assert d["omega_range"].tolist() == [0, 180] and d["eta_range"].tolist() == [0, 360]


def test_rejects_non_hdf5(tmp_path):
    from diffractom_amd.io import read_hdf5
    p = tmp_path / "x.h5"
    p.write_bytes(b"not an hdf5 file at all")
    with pytest.raises(ValueError):
        read_hdf5(p)
