"""The forward projection as a torch.autograd.Function (backward = the HIP adjoint) and the intensity layer
(differentiable/data_model.py:20-64), checked by the dot-product test, by the layer's explicit vjp, and by central
finite differences in FP64 of a Poisson + TV loss evaluated with the FP64 oracle's operator (64 random directions,
h = 1e-4; tolerance 1e-4 relative on the directional derivatives, the FP32 operator's own error being ~1e-5); the
finite-difference loss is built only from the oracle's pieces (operator, intensity layer, deviance, smoothed TV)."""
import numpy as np
import pytest

from pairs import build, cfg

pytestmark = pytest.mark.gpu


def _problem():
    c = cfg(nx=12, ny=12, my=13, n_omega=16, angle_range=(0, 180), cor=0.2, n_eta=24)
    op, orc, _ = build(c, K=6, sigma_deg=5.0, pf_mode="sparse")
    return op, orc


def test_dot_product_and_subset_rows():
    import torch
    from diffractom_amd.differentiable import forward_project
    op, _ = _problem()
    rng = np.random.default_rng(0)
    x = torch.from_numpy((rng.random(op.coeff_shape) * op.support_mask()[None]).astype(np.float32)).cuda()
    y = torch.from_numpy(rng.random(op.data_shape).astype(np.float32)).cuda()
    x.requires_grad_(True)
    u = forward_project(x, op)
    lhs = (u.double() * y.double()).sum()
    lhs.backward()
    aty = torch.from_numpy(op.adjoint(y.cpu().numpy())).cuda()
    assert torch.equal(x.grad, aty)
    rhs = (x.detach().double() * aty.double()).sum()
    assert abs(float(lhs) / float(rhs) - 1) < 1e-5
    # subset: rows of the other angles are 0 and the backward reads only the subset's rows
    subs = op.set_omega_subsets(4)
    x.grad = None
    u1 = forward_project(x, op, 1)
    rows = torch.from_numpy(subs[1]).cuda()
    mask = torch.ones(op.N_Omega, dtype=torch.bool, device="cuda")
    mask[rows] = False
    assert bool((u1[mask] == 0).all())
    assert torch.equal(u1[rows], u.detach()[rows])
    (u1 * y).sum().backward()
    ysub = torch.zeros_like(y)
    ysub[rows] = y[rows]
    assert torch.allclose(x.grad, torch.from_numpy(op.adjoint(ysub.cpu().numpy())).cuda(), rtol=1e-6, atol=1e-6)


def test_intensity_layer_matches_oracle():
    """The layer's forward and vjp (float64 and float32 parameters) against the FP64 oracle's independent
    restatement (oracle.model.intensity_forward / intensity_vjp, itself pinned by closed forms)."""
    import torch
    from diffractom_amd.differentiable import IntensityLayer
    from oracle.model import intensity_forward, intensity_vjp
    op, _ = _problem()
    lay = IntensityLayer(op.data_shape, op.N_peaks)
    rng = np.random.default_rng(5)
    prm_np = {"angle_scale": rng.uniform(0.9, 1.1, op.N_Omega), "trans_scale": rng.uniform(0.8, 1.2, op.My),
              "ring_scale": rng.uniform(0.5, 2.0, op.N_peaks), "background": rng.uniform(0, 3, op.N_seg)}
    u = rng.random(op.data_shape)
    r = rng.standard_normal(op.data_shape)
    args = (prm_np["angle_scale"], prm_np["trans_scale"], prm_np["ring_scale"])
    d_ref = intensity_forward(u, *args, prm_np["background"], op.N_peaks)
    g_ref = intensity_vjp(u, *args, r, op.N_peaks)
    for dt, tol in ((torch.float64, 1e-13), (torch.float32, 2e-6)):
        prm = {k: torch.from_numpy(v).to("cuda", dt) for k, v in prm_np.items()}
        d = lay.forward(torch.from_numpy(u).to("cuda", torch.float32 if dt == torch.float32 else dt), prm)
        assert np.abs(d.double().cpu().numpy() - d_ref).max() <= tol * np.abs(d_ref).max(), dt
        g = lay.vjp(torch.from_numpy(u).to("cuda", dt), prm, torch.from_numpy(r).to("cuda", dt))
        for k, v in g_ref.items():
            gv = g[k].double().cpu().numpy()
            assert np.abs(gv - v).max() <= tol * max(np.abs(v).max(), 1.0) * 50, (dt, k)


def test_intensity_layer_vjp_matches_autograd():
    import torch
    from diffractom_amd.differentiable import IntensityLayer
    op, _ = _problem()
    lay = IntensityLayer(op.data_shape, op.N_peaks)
    g = torch.Generator(device="cuda").manual_seed(3)
    prm = {k: (v + 0.1 * torch.rand(v.shape, generator=g, device="cuda", dtype=torch.float64)).requires_grad_(True)
           for k, v in lay.params0(device="cuda").items()}
    u = torch.rand(op.data_shape, generator=g, device="cuda", dtype=torch.float64).requires_grad_(True)
    r = torch.rand(op.data_shape, generator=g, device="cuda", dtype=torch.float64)
    (lay.forward(u, prm) * r).sum().backward()
    out = lay.vjp(u.detach(), {k: v.detach() for k, v in prm.items()}, r)
    for k in prm:
        assert torch.allclose(out[k], prm[k].grad, rtol=1e-12, atol=1e-12), k
    assert torch.allclose(out["u"], u.grad, rtol=1e-12, atol=1e-12)


def test_loss_gradients_vs_fp64_finite_differences():
    """L(c, f) = PoissonNLL(f_a * (A c); b) + 1e-3 TV(c): gradients with respect to c (autograd through the HIP
    operator) and the per-angle scales f_a ~ U(0.9, 1.1) against central differences of the same loss with the FP64
    oracle operator."""
    import torch
    from diffractom_amd.differentiable import IntensityLayer, forward_project, poisson_nll, total_variation
    from oracle.model import deviance, intensity_forward, smooth_tv
    op, orc = _problem()
    rng = np.random.default_rng(7)
    supp = op.support_mask()[None]
    c0 = (0.2 + rng.random(op.coeff_shape)) * supp
    f0 = rng.uniform(0.9, 1.1, op.N_Omega)
    b = rng.poisson(np.maximum(orc.direct(c0), 0) * 20).astype(np.float64)
    lay = IntensityLayer(op.data_shape, op.N_peaks)
    bt = torch.from_numpy(b)
    ones = lambda n: np.ones(n)

    def loss_fp64(c, f):
        # every piece from the FP64 oracle: operator, intensity layer, Poisson deviance, smoothed TV
        d = intensity_forward(orc.direct(c), f, ones(op.My), ones(op.N_peaks), np.zeros(op.N_seg), op.N_peaks)
        return deviance(d, b) + 1e-3 * smooth_tv(c)

    # the product's loss pieces equal the oracle's at the expansion point
    prm64 = lay.params0()
    prm64["angle_scale"] = torch.from_numpy(f0)
    u0 = torch.from_numpy(orc.direct(c0))
    ref0 = loss_fp64(c0, f0)
    own0 = float(poisson_nll(lay.forward(u0, prm64), bt) + 1e-3 * total_variation(torch.from_numpy(c0)))
    assert abs(own0 / ref0 - 1) < 1e-9, (own0, ref0)        # (the oracle takes eps in float32, as the kernels)

    # analytic gradients on the GPU
    c = torch.from_numpy(c0.astype(np.float32)).cuda().requires_grad_(True)
    prm = lay.params0(device="cuda")
    prm["angle_scale"] = torch.from_numpy(f0).cuda().requires_grad_(True)
    L = poisson_nll(lay.forward(forward_project(c, op), prm), bt.cuda()) + 1e-3 * total_variation(c)
    L.backward()
    gc = c.grad.double().cpu().numpy()
    gf = prm["angle_scale"].grad.double().cpu().numpy()
    h = 1e-4
    ad, fd = [], []
    for i in range(64):
        v = rng.standard_normal(op.coeff_shape) * supp
        w = rng.standard_normal(op.N_Omega)
        fd.append((loss_fp64(c0 + h * v, f0 + h * w) - loss_fp64(c0 - h * v, f0 - h * w)) / (2 * h))
        ad.append(float((gc * v).sum() + (gf * w).sum()))
    ad, fd = np.array(ad), np.array(fd)
    err = np.linalg.norm(ad - fd) / np.linalg.norm(fd)
    print(f"directional derivatives: relative L2 {err:.2e}", flush=True)
    assert err < 1e-4


def test_total_variation_kernel_matches_float64():
    """The HIP smooth-TV pass (float32 device coefficients: value with float64 partial sums, gradient in the same
    pass) against the float64 torch formula and its autograd gradient on the CPU."""
    import torch
    from diffractom_amd.differentiable import total_variation
    g = torch.Generator().manual_seed(2)
    c = torch.rand((3, 9, 11), generator=g, dtype=torch.float64)
    c[1, 2:5, 3:7] = 0.25                                  # flat patches: the eps branch
    cd = c.float().cuda().requires_grad_(True)
    v = total_variation(cd)
    v.backward()
    cc = c.float().double().requires_grad_(True)
    v64 = total_variation(cc)
    v64.backward()
    assert abs(float(v) / float(v64) - 1) < 1e-6
    err = (cd.grad.double().cpu() - cc.grad).abs().max() / cc.grad.abs().max()
    assert float(err) < 1e-5, float(err)
