"""The allocation-free deviance (dfm_deviance_noalloc: partial sums in the library's own device scratch, summed on
the device in index order) equals the deviance of dfm_deviance_partial (caller-allocated partial sums, added in
index order on the host) bit for bit, on two calls in a row (the second with a smaller input, so the scratch still
holds the first call's partial sums beyond the second's), with weights, gain and background, and also when the
second call has one partial sum. The FP64 oracle's deviance of the same inputs agrees within 1e-12."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _inputs(n, seed):
    import torch
    rng = np.random.default_rng(seed)
    mu = rng.uniform(-0.5, 50.0, n).astype(np.float32)
    b = rng.poisson(np.maximum(mu, 0) + 1).astype(np.float32)
    w = (rng.random(7) > 0.3).astype(np.float32)                     # periodic weights, period 7
    g = rng.uniform(0.5, 2.0, 5).astype(np.float32)                  # periodic gain, period 5
    bg = rng.uniform(0.0, 1.0, 3).astype(np.float32)                 # periodic background, period 3
    dev = {k: torch.from_numpy(v).cuda() for k, v in dict(mu=mu, b=b, w=w, g=g, bg=bg).items()}
    return dev, dict(mu=mu, b=b, w=w, g=g, bg=bg)


def _with_partials(d, n, eps):
    import torch
    from diffractom_amd import _lib
    part = torch.empty(_lib.MAX_PARTIALS, dtype=torch.float64, device="cuda")
    npart = ctypes.c_int32(0)
    _lib.call("dfm_deviance_partial", _lib.ptr(d["mu"]), _lib.ptr(d["b"]), _lib.ptr(d["w"]), 7, _lib.ptr(d["g"]), 5,
              _lib.ptr(d["bg"]), 3, eps, n, _lib.ptr(part), ctypes.byref(npart), None)
    torch.cuda.synchronize()
    s = 0.0
    for v in part[: npart.value].cpu().numpy().tolist():        # index order, as the device sum
        s += v
    return s, npart.value


def _noalloc(d, n, eps):
    import torch
    from diffractom_amd import _lib
    out = torch.full((1,), float("nan"), dtype=torch.float64, device="cuda")
    _lib.call("dfm_deviance_noalloc", _lib.ptr(d["mu"]), _lib.ptr(d["b"]), _lib.ptr(d["w"]), 7, _lib.ptr(d["g"]), 5,
              _lib.ptr(d["bg"]), 3, eps, n, _lib.ptr(out), None)
    torch.cuda.synchronize()
    return float(out.item())


def test_noalloc_deviance_matches_partials_twice():
    from oracle.model import deviance, periodic
    eps = 1e-3
    for n, seed in ((600_000, 1), (40_000, 2), (200, 3)):
        print(f"case noalloc deviance n={n}", flush=True)
        d, h = _inputs(n, seed)
        ref, npart = _with_partials(d, n, eps)
        got = _noalloc(d, n, eps)
        assert got == ref, (n, npart, got, ref)
        orc = deviance(h["mu"].astype(np.float64), h["b"].astype(np.float64), periodic(h["w"], n),
                       periodic(h["g"], n), periodic(h["bg"], n), float(np.float32(eps)))
        assert abs(got - orc) <= 1e-12 * abs(orc), (got, orc)


def test_noalloc_rejects_bad_arguments():
    from diffractom_amd import _lib
    with pytest.raises(ValueError, match="bad arguments"):
        _lib.call("dfm_deviance_noalloc", None, None, None, 0, None, 0, None, 0, 1e-3, 10, None, None)
