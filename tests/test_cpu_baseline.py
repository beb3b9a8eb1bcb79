"""The host-core baseline's CSR products (oracle/c/cpu_step.c, timed by bench.py as the CPU baseline) against dense
NumPy products on a random sparse PF whose orientation count (37) is below the sinogram's row stride (40): the
adjoint CSR has R * Kb rows, the sinogram columns from Kb up are written as 0 (CPU)."""
import ctypes

import numpy as np

from oracle import cpu_baseline


def _p(a):
    return ctypes.c_void_p(a.ctypes.data)


def test_csr_products_match_dense():
    lib = cpu_baseline._load()
    lib.cpu_set_threads(2)
    rng = np.random.default_rng(0)
    R, kb, K, cp, my = 3, 37, 40, 12, 5
    pf = rng.random((R, kb, cp)) * (rng.random((R, kb, cp)) < 0.3)
    pf = pf.astype(np.float32)
    # forward rows (r, j) listing k ascending; adjoint rows (r, k) listing j ascending
    rpf, ck, vf = [0], [], []
    for r in range(R):
        for j in range(cp):
            k = np.nonzero(pf[r, :, j])[0]
            ck += k.tolist()
            vf += pf[r, k, j].tolist()
            rpf.append(len(ck))
    rpa, cj, va = [0], [], []
    for r in range(R):
        for k in range(kb):
            j = np.nonzero(pf[r, k])[0]
            cj += j.tolist()
            va += pf[r, k, j].tolist()
            rpa.append(len(cj))
    rpf, rpa = np.array(rpf, np.int32), np.array(rpa, np.int32)
    ck, cj = np.array(ck, np.uint16), np.array(cj, np.uint16)
    vf, va = np.array(vf, np.float32), np.array(va, np.float32)
    om = np.array([2, 0], np.int32)
    sino = rng.random((len(om), my, K)).astype(np.float32)
    data = np.full((R, my, cp), -1.0, np.float32)
    lib.cpu_spmm_forward(_p(sino), _p(rpf), _p(ck), _p(vf), _p(data), _p(om), len(om), my, cp, K)
    for a, r in enumerate(om):
        np.testing.assert_allclose(data[r], sino[a, :, :kb] @ pf[r], rtol=1e-5)
    assert np.all(data[1] == -1.0)
    out = np.full((len(om), my, K), 5.0, np.float32)
    lib.cpu_spmm_adjoint(_p(data), _p(rpa), _p(cj), _p(va), _p(out), _p(om), len(om), my, cp, K, kb,
                         ctypes.c_float(0.5))
    for a, r in enumerate(om):
        np.testing.assert_allclose(out[a, :, :kb], 0.5 * data[r] @ pf[r].T, rtol=1e-5)
        assert np.all(out[a, :, kb:] == 0.0)
