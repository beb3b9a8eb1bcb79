"""Gradients with respect to the model's parameters on the GPU (PFGradient, ProjGradient) against the FP64
differentiable model (oracle/diffmodel.py, itself pinned by finite differences in test_oracle_diffmodel.py), at the
reference's bar of 1e-4 of the largest component (tests/test_differentiable.py:69-111), for several PF storages
and precisions (the gradient evaluates the PF's derivatives itself), with several orientation batches; the
Gauss-Newton blocks; and the in-place updates: after set_pf_parameters / set_proj_parameters the operator's forward
equals the model's at the new parameters (2e-5 of the largest entry), set_geometry is absolute."""
import numpy as np
import pytest

from oracle.diffmodel import PF_GROUPS, OracleDiffModel
from pairs import build, cfg, rel_max

pytestmark = pytest.mark.gpu


def _setup(pf_mode="sparse", precision="fp32", kernel="trapezoid", quadrature="auto", seed=0):
    c = cfg(nx=14, ny=12, my=17, n_omega=16, angle_range=(0, 360), cor=0.7, n_eta=24)
    op, orc, _ = build(c, K=12, sigma_deg=6.0, seed=5, pf_mode=pf_mode, pf_precision=precision, max_gb=1e-5,
                       radon_kernel=kernel, quadrature=quadrature)
    rng = np.random.default_rng(seed)
    x = (rng.random(op.coeff_shape) * op.support_mask()[None]).astype(np.float32)
    r = rng.normal(size=op.data_shape).astype(np.float32)
    return op, orc, x, r, rng


@pytest.mark.parametrize("pf_mode,precision", [("sparse", "fp32"), ("generated", "fp16"), ("dense", "bf16_residual")])
def test_pf_gradients_match_the_model(pf_mode, precision):
    from diffractom_amd.differentiable import PFGradient
    print(f"case pf gradients {pf_mode} {precision}", flush=True)
    op, orc, x, r, _ = _setup(pf_mode, precision)
    assert len(op.batches) > 1
    m = OracleDiffModel(orc)
    g_ref = m.vjp(x, m.params0(), r, groups=PF_GROUPS)
    g = PFGradient(op)(x, r)
    for k in PF_GROUPS:
        a, b = np.atleast_1d(g_ref[k]), np.atleast_1d(g[k])
        err = np.abs(a - b).max() / np.abs(a).max()
        print(f"  {k}: {err:.2e}", flush=True)
        assert err < 1e-4, k


def test_gauss_newton_blocks():
    import torch
    from diffractom_amd.differentiable import PFGradient
    print("case gauss-newton blocks", flush=True)
    op, orc, x, r, rng = _setup()
    curv = rng.random(op.data_shape).astype(np.float32)
    m = OracleDiffModel(orc)
    prm = m.params0()
    Pm, _ = m.projection(prm)
    S = np.einsum("asp,kp->ask", Pm, x.reshape(op.K, -1).astype(np.float64))
    H_ref = m.pf_jacobian_blocks(prm, curv, S)
    out = PFGradient(op, hessian=True)(torch.from_numpy(x).cuda(), torch.from_numpy(r).cuda(), curv=curv)
    assert rel_max(out["hess"], H_ref) < 1e-4
    g_ref = m.vjp(x, prm, r, groups=("rot", "sigma"))
    assert rel_max(out["rot"], g_ref["rot"]) < 1e-4 and rel_max(out["sigma"], g_ref["sigma"]) < 1e-4


@pytest.mark.parametrize("kernel", ["trapezoid", "hat"])
def test_projector_gradients(kernel):
    from diffractom_amd.differentiable import ProjGradient
    print(f"case projector gradients {kernel}", flush=True)
    op, orc, x, r, _ = _setup(kernel=kernel, quadrature="midpoint")
    m = OracleDiffModel(orc)
    g_ref = m.vjp(x, m.params0(), r, groups=("shift", "cor"))
    g = ProjGradient(op)(x, r)
    assert rel_max(g["shift"], g_ref["shift"]) < 1e-4
    assert abs(g["cor"] - g_ref["cor"]) < 1e-4 * abs(g_ref["cor"])


@pytest.mark.parametrize("pf_mode,precision", [("sparse", "fp32"), ("generated", "bf16_residual")])
def test_set_pf_parameters_moves_the_operator(pf_mode, precision):
    from diffractom_amd.differentiable import rotate_orientations, set_pf_parameters
    print(f"case set_pf_parameters {pf_mode} {precision}", flush=True)
    op, orc, x, _, rng = _setup(pf_mode, precision)
    op.set_omega_subsets(4)
    m = OracleDiffModel(orc)
    prm = m.params0()
    prm["rot"] = rng.normal(0, 0.02, (op.K, 3))
    prm["sigma"] = (prm["sigma"] * 1.1).astype(np.float32).astype(np.float64)        # the operator keeps float32
    prm["two_theta"] = (prm["two_theta"] + 0.001).astype(np.float32).astype(np.float64)
    prm["eta_offset"] = 0.01
    prm["axis_tilt"] = np.array([0.01, -0.02])
    d_ref = m.forward(x, prm)
    set_pf_parameters(op, rotations=rotate_orientations(op.grid_rot_cpu, prm["rot"]), sigma=prm["sigma"],
                      two_theta=prm["two_theta"], eta_offset=prm["eta_offset"], axis_tilt=prm["axis_tilt"])
    tol = 2e-5 if precision == "fp32" else 3e-5
    assert rel_max(op.direct(x), d_ref) < tol
    # the subsets were set up again on the new PF
    import torch
    d = torch.zeros(op.data_shape, device="cuda")
    op.direct_cl(torch.from_numpy(x).cuda(), d, subset=1)
    rows = op.omega_subsets[1]
    assert rel_max(d.cpu().numpy()[rows], d_ref[rows]) < tol


def test_set_proj_parameters_and_geometry():
    from diffractom_amd.differentiable import set_geometry, set_proj_parameters
    print("case set_proj_parameters / set_geometry", flush=True)
    op, orc, x, _, rng = _setup()
    d0 = op.direct(x)
    m = OracleDiffModel(orc)
    prm = m.params0()
    prm["cor"] = prm["cor"] + 0.3
    prm["shift"] = rng.normal(0, 0.4, op.N_Omega)
    set_proj_parameters(op, cor=prm["cor"], shift=prm["shift"])
    assert rel_max(op.direct(x), m.forward(x, prm)) < 2e-5
    # absolute: the same eta offset twice is the same operator, and zero offsets return to the initial one
    set_geometry(op, eta_offset=0.02)
    d1 = op.direct(x)
    set_geometry(op, eta_offset=0.02)
    assert np.array_equal(op.direct(x), d1)
    set_geometry(op, eta_offset=0.0, cor=orc.cor, shift=np.zeros(op.N_Omega))
    assert rel_max(op.direct(x), d0) < 1e-6
