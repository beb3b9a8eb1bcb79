"""The Fourier-slice projector (projector="fft") against the FP64 Fourier-slice model (oracle/fourier.py, pinned by
a closed form in test_oracle_fourier.py): projections within 2e-5 of the largest entry (the non-uniform FFT's
~1e-5 plus float32), the backward projection the weighted transpose of the forward (1e-6), several channel slices;
the operator with projector="fft": forward / adjoint against the oracle PF with the model's projections, subsets,
and the SPDHG backprojection update."""
import numpy as np
import pytest

from oracle.fourier import OracleFourierRadon
from pairs import build, cfg, rel_max

pytestmark = pytest.mark.gpu


def _op(K=9, **kw):
    c = cfg(nx=20, ny=17, my=27, n_omega=14, angle_range=(0, 180), cor=0.6, n_eta=12)
    return build(c, K=K, pf_mode="sparse", projector="fft", **kw)


def _model(op):
    return OracleFourierRadon(op.Nx, op.Ny, op.angles, op.My, detector_shift=op.cor_offset)


def test_projector_matches_the_model_and_is_its_transpose():
    import torch
    print("case fourier projector", flush=True)
    op, _, _ = _op(oracle=False)
    fr = _model(op)
    assert op.fourier.L == fr.L
    op.fourier.max_channels = 4                       # several channel slices
    rng = np.random.default_rng(0)
    C = 11
    img = rng.random((C, op.Ny, op.Nx))
    img_k = torch.from_numpy(np.ascontiguousarray(img.reshape(C, -1).T, dtype=np.float32)).cuda()   # (Npix, C)
    angles = torch.arange(op.N_Omega, dtype=torch.int32, device="cuda")
    sino = torch.empty(op.N_Omega * op.My * C, device="cuda")
    st = op._stream()
    op.fourier.forward(img_k, C, 0, C, sino, angles, st)
    got = sino.view(op.N_Omega, op.My, C).cpu().numpy()
    assert rel_max(got, fr.forward(img)) < 2e-5
    y = rng.random((op.N_Omega, op.My, C)).astype(np.float32)
    out = torch.zeros(op.Nx * op.Ny * C, device="cuda")
    op.fourier.backward(torch.from_numpy(y).cuda().view(-1), angles, out, C, 0, C, st)
    back = out.view(op.Ny * op.Nx, C).cpu().numpy().T.reshape(C, op.Ny, op.Nx)
    w = np.pi / op.N_Omega / op.radon_scale
    assert rel_max(back, fr.backward(y, w)) < 2e-5
    lhs = float(np.sum(got.astype(np.float64) * y) * w)
    rhs = float(np.sum(img * back))
    assert abs(lhs / rhs - 1) < 1e-6
    # a subset of angles: only its rows, in its order
    sub = torch.tensor([3, 9, 11], dtype=torch.int32, device="cuda")
    s2 = torch.empty(3 * op.My * C, device="cuda")
    op.fourier.forward(img_k, C, 0, C, s2, sub, st)
    assert rel_max(s2.view(3, op.My, C).cpu().numpy(), got[[3, 9, 11]]) < 1e-6


def test_operator_with_the_fourier_projector():
    import torch
    print("case operator projector=fft", flush=True)
    op, orc, _ = _op()
    assert op.projector == "fft" and op.radon_scale == 1.0
    fr = _model(op)
    rng = np.random.default_rng(1)
    x = (rng.random(op.coeff_shape) * op.support_mask()[None]).astype(np.float32)
    y = rng.random(op.data_shape).astype(np.float32)
    S = fr.forward(x)                                               # (R, Ns, K)
    d_ref = np.stack([(orc.pf_sparse(a).T @ S[a].T).T for a in range(op.N_Omega)])
    assert rel_max(op.direct(x), d_ref) < 3e-5
    Z = np.stack([(orc.pf_sparse(a) @ y[a].astype(np.float64).T).T for a in range(op.N_Omega)])   # (R, Ns, K)
    a_ref = fr.backward(Z)
    assert rel_max(op.adjoint(y), a_ref) < 3e-5
    subs = op.set_omega_subsets(3)
    xt, yt = torch.from_numpy(x).cuda(), torch.from_numpy(y).cuda()
    full = op.direct(xt)
    d = torch.full(op.data_shape, -7.0, device="cuda")
    op.direct_cl(xt, d, subset=1)
    rows = torch.from_numpy(subs[1]).cuda()
    assert rel_max(d[rows].cpu().numpy(), full[rows].cpu().numpy()) < 1e-6
    acc = torch.zeros(op.coeff_shape, device="cuda")
    part = torch.empty_like(acc)
    for m in range(3):
        op.adjoint_cl(yt, part, subset=m)
        acc += part
    assert rel_max(acc.cpu().numpy(), op.adjoint(yt).cpu().numpy()) < 1e-5
    # the SPDHG backprojection update (mode 1) = the plain backprojection and the update formula
    n_ang, _, _, geo_b = op._angle_set(1)
    b = op.batches[0]
    npix = op.Nx * op.Ny
    sino = torch.rand(n_ang * op.My * b["kstride"], device="cuda")
    img = torch.zeros(npix * op.K_stride, device="cuda")
    op._radon_backward(sino, geo_b, None, n_ang, img, op.K_stride, b["koff"], b["kstride"])
    x0 = torch.rand(npix * op.K_stride, device="cuda")
    z0 = torch.rand(npix * op.K_stride, device="cuda") - 0.5
    T = torch.rand(npix * op.K_stride, device="cuda") * (torch.rand(npix * op.K_stride, device="cuda") > 0.2)
    x1, z1 = x0.clone(), z0.clone()
    op._radon_backward(sino, geo_b, None, n_ang, None, op.K_stride, b["koff"], b["kstride"], x=x1, z=z1, T=T,
                       inv_p=0.7, lam=0.05)
    zn = z0 + img
    xn = torch.where(T > 0, torch.clamp(x0 - T * ((zn + 0.7 * img) + 0.05), min=0.0), torch.zeros_like(x0))
    assert rel_max(z1.cpu().numpy(), zn.cpu().numpy()) < 1e-6
    assert rel_max(x1.cpu().numpy(), xn.cpu().numpy()) < 1e-6


def test_spdhg_runs_on_the_fourier_projector():
    from diffractom_amd import SPDHGPoisson
    print("case spdhg projector=fft", flush=True)
    op, orc, _ = _op(K=12)
    rng = np.random.default_rng(2)
    xt = (rng.random(op.coeff_shape) * op.support_mask()[None]).astype(np.float32)
    b = rng.poisson(np.maximum(op.direct(xt), 0) * 50 / max(op.direct(xt).mean(), 1e-9)).astype(np.float32)
    sol = SPDHGPoisson(op, subsets=4, seed=0)
    sol.run(np.zeros(op.coeff_shape, np.float32), b, 6, diagnostics_interval=2)
    f = [s["f"] for s in sol.iter_stats if s.get("f") is not None]
    assert len(f) >= 2 and f[-1] < f[0]
