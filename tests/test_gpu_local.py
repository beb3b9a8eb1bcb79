"""The local-basis operator and SPDHG on it (operators/local_forward_operator.py, spdhg_poisson.py:754-943) on the
GPU, against the FP64 oracle (oracle/local.py: the global oracle on the scattered coefficients) and against the
global operator. The problems follow the reference's tests/test_subsets.py:174-259 (every orientation in one or two
discs of a 20 x 20 grid, 30 angles, several orientation batches) and tests/test_spdhg_tv.py:172-185.
Tolerances: FP32 forward / adjoint within 2e-5 of the largest entry (the operator's own bar), the matrix-core
precisions at their operator bars (bf16_residual 3e-5, fp16 2e-3), adjointness 1e-5, SPDHG iterates 1e-4 (relative
L2) and objectives 1e-5 against FP64."""
import numpy as np
import pytest

from oracle.local import OracleLocalOperator
from oracle.model import OracleOperator, OracleSPDHG
from pairs import AL3, build, cfg, orientations, rel_l2, rel_max

pytestmark = pytest.mark.gpu

TOL = {"fp32": 2e-5, "bf16_residual": 3e-5, "fp16": 2e-3}


def _discs(K, n, seed=3):
    rng = np.random.default_rng(seed)
    yy, xx = np.mgrid[:n, :n]
    mask = np.zeros((K, n, n), bool)
    for k in range(K):
        for _ in range(1 + k % 2):
            cy, cx, r = rng.uniform(3, n - 3), rng.uniform(3, n - 3), rng.uniform(1.5, 4)
            mask[k] |= (yy - cy) ** 2 + (xx - cx) ** 2 <= r * r
    return mask


def _local(pf_mode="sparse", split=True, K=30, max_gb=1e-4, full=False, precision="fp32", c=None, **kw):
    """(local operator, local oracle, global oracle, U) on the reference's local-basis test geometry."""
    from diffractom_amd import LocalSinglePhaseForwardOperator, LocalSupport, Material, MatrixGrid
    c = cfg(nx=20, ny=20, my=21, n_omega=30, angle_range=(0, 360), cor=0.5) if c is None else c
    U = orientations(K, 5)
    sigma = np.deg2rad(4.0)
    mat = Material.from_lattice_parameters(a=4.05, b=4.05, c=4.05, alpha=90, beta=90, gamma=90,
                                           symmetry_group="cubic", wavelength_A=0.35, hkl_list=AL3)
    mask = np.ones((K, c["Ny"], c["Nx"]), bool) if full else _discs(K, c["Nx"])
    op = LocalSinglePhaseForwardOperator(c, mat, MatrixGrid(U, sigma), max_gb, support=LocalSupport.from_mask(mask),
                                         normalized=True, pf_mode=pf_mode, split_patches=split,
                                         pf_precision=precision, **kw)
    base = OracleOperator(c, (4.05, 4.05, 4.05, 90, 90, 90), "cubic", AL3, 0.35, U, sigma, orthonormalize=False)
    k, iy, ix, pi = op._pairs()
    return op, OracleLocalOperator(base, k, iy, ix, pi, op.n_coef), base, U


@pytest.mark.parametrize("pf_mode,precision,staging", [("sparse", "fp32", True), ("sparse", "fp32", False),
                                                       ("generated", "fp32", True), ("sparse", "bf16_residual", True),
                                                       ("generated", "fp16", True)])
@pytest.mark.parametrize("split", [True, False])
def test_local_matches_oracle(pf_mode, precision, staging, split):
    """staging False: the global-memory forms of the strip projection and the backprojection (taken by patches and
    strips too large for the LDS) and the counted pixel loop (taken by pixels smaller than about 0.4 bins)."""
    print(f"case pf_mode={pf_mode} precision={precision} staging={staging} split={split}", flush=True)
    op, orc, _, _ = _local(pf_mode, split, precision=precision)
    op.local_staging = staging
    if not staging:
        op._n_reach = 0
    assert len(op.batches) > 1 and op.n_patches >= op.K
    rng = np.random.default_rng(6)
    x = rng.random(op.coeff_shape).astype(np.float32)
    y = rng.random(op.data_shape).astype(np.float32)
    d, d_ref = op.direct(x), orc.direct(x)
    a, a_ref = op.adjoint(y), orc.adjoint(y)
    print(f"  forward {rel_max(d, d_ref):.2e} adjoint {rel_max(a, a_ref):.2e}", flush=True)
    assert rel_max(d, d_ref) < TOL[precision]
    assert rel_max(a, a_ref) < TOL[precision]
    mask = op.coefficient_mask()
    assert np.all(a[mask == 0] == 0.0)
    lhs = float(np.dot(x.astype(np.float64), a))
    rhs = float(np.dot(d.ravel().astype(np.float64), y.ravel()))
    assert abs(lhs / rhs - 1) < (1e-5 if precision != "fp16" else 1e-3)


@pytest.mark.parametrize("kernel", ["trapezoid", "hat"])
def test_local_equals_global_operator(kernel):
    """Every orientation on the whole field of view, one patch each: the local operator gives the global operator's
    forward of the same (zero-outside-the-view) coefficients and its adjoint on the view."""
    print(f"case full support, kernel {kernel}", flush=True)
    op, _, _, U = _local("sparse", split=False, full=True, radon_kernel=kernel)
    gop, _, _ = build(cfg(nx=20, ny=20, my=21, n_omega=30, angle_range=(0, 360), cor=0.5), K=30, pf_mode="sparse",
                      max_gb=1e-4, oracle=False, radon_kernel=kernel)
    fov = gop.support_mask()
    rng = np.random.default_rng(2)
    xd = (rng.random(gop.coeff_shape) * fov[None]).astype(np.float32)
    y = rng.random(gop.data_shape).astype(np.float32)
    assert rel_max(op.direct(op.from_dense(xd)), gop.direct(xd)) < 1e-5
    assert rel_max(op.to_dense(op.adjoint(y)), gop.adjoint(y) * fov[None]) < 1e-5
    assert op.n_coef >= int(fov.sum()) * 30


@pytest.mark.parametrize("pf_mode", ["sparse", "generated"])
def test_local_subsets(pf_mode):
    """tests/test_subsets.py:197-231: the subset forward gives the full forward's rows (other rows untouched), the
    subset adjoint the adjoint of the data zeroed outside the subset's rows, the subset adjoints of a partition sum
    to the full adjoint, each subset adjoint is the transpose of its forward."""
    import torch
    print(f"case pf_mode={pf_mode}", flush=True)
    op, _, _, _ = _local(pf_mode)
    R = op.N_Omega
    rng = np.random.default_rng(6)
    x = rng.random(op.coeff_shape).astype(np.float32)
    y = rng.random(op.data_shape).astype(np.float32)
    full, adj = op.direct(x), op.adjoint(y)
    for subsets in (4, [np.arange(0, 7), np.arange(7, 20), np.arange(20, R)]):
        subs = op.set_omega_subsets(subsets)
        xd, yd = torch.from_numpy(x).cuda(), torch.from_numpy(y).cuda()
        adj_sum = np.zeros_like(adj, dtype=np.float64)
        for m, s in enumerate(subs):
            d = torch.full(op.data_shape, -7.0, dtype=torch.float32, device="cuda")
            op.direct_cl(xd, d, subset=m)
            d = d.cpu().numpy()
            assert rel_max(d[s], full[s]) < 1e-6
            assert np.all(d[np.setdiff1d(np.arange(R), s)] == -7.0)
            a_m = torch.empty(op.coeff_shape, dtype=torch.float32, device="cuda")
            op.adjoint_cl(yd, a_m, subset=m)
            a_m = a_m.cpu().numpy()
            ym = np.zeros_like(y)
            ym[s] = y[s]
            assert rel_max(a_m, op.adjoint(ym)) < 1e-6
            adj_sum += a_m
            lhs = float(np.dot(d[s].ravel().astype(np.float64), y[s].ravel()))
            rhs = float(np.dot(x.astype(np.float64), a_m))
            assert abs(lhs / rhs - 1) < 1e-5
        assert rel_max(adj_sum, adj) < 1e-5


def _counts(orc, mask, seed=7):
    rng = np.random.default_rng(seed)
    xt = rng.random(orc.coeff_shape) * (rng.random(orc.coeff_shape) < 0.3) * mask
    ax = orc.direct(xt)
    return rng.poisson(ax / ax.mean() * 50).astype(np.float32)


@pytest.mark.parametrize("tv", [0.0, 1.0])
def test_local_spdhg_matches_oracle(tv):
    """SPDHG on the local basis = SPDHG on (K, Ny, Nx) with the coefficients outside the support held at zero (one
    patch per orientation, so the total variation's neighbours agree too): iterate, objectives, dual state."""
    from diffractom_amd import SPDHGPoisson
    print(f"case tv={tv}", flush=True)
    op, orc, base, _ = _local("sparse", split=False)
    mask = op.coefficient_mask()
    b = _counts(orc, mask)
    sp = SPDHGPoisson(op, subsets=4, seed=1, tv=tv)
    x = sp.run(np.zeros(op.coeff_shape, np.float32), b, 8, diagnostics_interval=2)
    ref = OracleSPDHG(base, subsets=4, seed=1, tv=tv, support=orc.support_mask3())
    x_ref = ref.run(np.zeros(base.coeff_shape), b, 8, diagnostics_interval=2)
    assert abs(sp.rho_used / ref.rho_used - 1) < 1e-6
    print(f"  x {rel_l2(op.to_dense(x), x_ref):.2e}", flush=True)
    assert rel_l2(op.to_dense(x), x_ref) < 1e-4
    f = [s["f"] for s in sp.iter_stats if s["f"] is not None]
    f_ref = [s["f"] for s in ref.iter_stats if s["f"] is not None]
    np.testing.assert_allclose(f, f_ref, rtol=1e-5)
    # the dual state on the data points the support reaches (a row sum A 1 at float rounding level makes the step
    # sigma = gamma / (A 1) arbitrary: such points flip between 0 and the limit of the prox in either arithmetic)
    rows = base.direct(orc.support_mask3().astype(np.float64))
    ok = rows > 1e-4 * rows.max()
    assert rel_l2(sp.state["y"][ok], ref.state["y"][ok]) < 1e-4
    assert rel_l2(sp.state["z"], orc.gather(ref.state["z"])) < 1e-4


def test_local_spdhg_split_patches_matches_oracle():
    """Several patches per orientation (no total variation): the same iterates as SPDHG with the support mask."""
    from diffractom_amd import SPDHGPoisson
    op, orc, base, _ = _local("generated", split=True)
    assert op.n_patches > op.K
    b = _counts(orc, op.coefficient_mask(), seed=8)
    sp = SPDHGPoisson(op, subsets=3, seed=2)
    x = sp.run(np.zeros(op.coeff_shape, np.float32), b, 6, diagnostics_interval=3)
    ref = OracleSPDHG(base, subsets=3, seed=2, support=orc.support_mask3())
    x_ref = ref.run(np.zeros(base.coeff_shape), b, 6, diagnostics_interval=3)
    assert rel_l2(op.to_dense(x), x_ref) < 1e-4
    np.testing.assert_allclose([s["f"] for s in sp.iter_stats if s["f"] is not None],
                               [s["f"] for s in ref.iter_stats if s["f"] is not None], rtol=1e-5)


def test_local_spdhg_full_support_equals_dense():
    """tests/test_spdhg_tv.py:172-185: the local basis with every orientation on the whole view (one patch each)
    gives the dense run's iterate, total variation included."""
    from diffractom_amd import SPDHGPoisson
    op, orc, base, _ = _local("sparse", split=False, full=True)
    gop, _, _ = build(cfg(nx=20, ny=20, my=21, n_omega=30, angle_range=(0, 360), cor=0.5), K=30, pf_mode="sparse",
                      max_gb=1e-4, oracle=False)
    b = _counts(orc, op.coefficient_mask(), seed=9)
    xs = {}
    for name, o in (("dense", gop), ("local", op)):
        sp = SPDHGPoisson(o, subsets=3, tv=1.5, seed=4)
        x = sp.run(np.zeros(o.coeff_shape, np.float32), b, 10, diagnostics_interval=10)
        xs[name] = x if name == "dense" else op.to_dense(x)
    np.testing.assert_allclose(xs["local"], xs["dense"], rtol=1e-4, atol=1e-5 * np.abs(xs["dense"]).max())


def test_local_spdhg_resume_and_graphs():
    """A run continued from solver.state equals one run (z kept; a state without z recomputes z = A^T y), and the
    HIP-graph replay of the steps gives the eager iterates."""
    from diffractom_amd import SPDHGPoisson
    op, orc, _, _ = _local("sparse", split=True)
    b = _counts(orc, op.coefficient_mask(), seed=10)
    one = SPDHGPoisson(op, subsets=3, seed=1)
    x_one = one.run(np.zeros(op.coeff_shape, np.float32), b, 6, diagnostics_interval=2)
    two = SPDHGPoisson(op, subsets=3, seed=1)
    x = two.run(np.zeros(op.coeff_shape, np.float32), b, 4, diagnostics_interval=4)
    x = two.run(x, b, 2, diagnostics_interval=2, state=two.state)
    assert two.state["epoch"] == 6
    np.testing.assert_allclose(x, x_one, rtol=1e-5, atol=1e-6 * np.abs(x_one).max())
    z = op.adjoint(two.state["y"])
    np.testing.assert_allclose(two.state["z"], z, rtol=1e-4, atol=1e-5 * np.abs(z).max())
    g = SPDHGPoisson(op, subsets=3, seed=1)
    g.graphs = True
    x_g = g.run(np.zeros(op.coeff_shape, np.float32), b, 6, diagnostics_interval=2)
    np.testing.assert_allclose(x_g, x_one, rtol=1e-6, atol=1e-7 * np.abs(x_one).max())


def test_local_fista_equals_dense():
    """FISTAPoisson runs on the local basis (the support held by the operator) and, with every orientation on the
    whole view, gives the dense run's iterate."""
    from diffractom_amd import FISTAPoisson
    op, orc, _, _ = _local("sparse", split=False, full=True)
    gop, _, _ = build(cfg(nx=20, ny=20, my=21, n_omega=30, angle_range=(0, 360), cor=0.5), K=30, pf_mode="sparse",
                      max_gb=1e-4, oracle=False)
    b = _counts(orc, op.coefficient_mask(), seed=11)
    xd = FISTAPoisson(gop).run(np.zeros(gop.coeff_shape, np.float32), b, 15)
    xl = FISTAPoisson(op).run(np.zeros(op.coeff_shape, np.float32), b, 15)
    assert rel_l2(op.to_dense(xl), xd) < 1e-4


def test_local_rejects_other_layout_solvers():
    from diffractom_amd import OSEMPoisson, PDHGPoisson
    op, _, _, _ = _local("sparse", split=False, K=8)
    with pytest.raises(NotImplementedError):
        OSEMPoisson(op)
    with pytest.raises(NotImplementedError):
        PDHGPoisson(op)
