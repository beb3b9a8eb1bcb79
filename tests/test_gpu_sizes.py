"""The measured configurations at their real sizes (the small-shape parity tests pin the kernels against the FP64
oracle; these check that the same paths hold at the benchmark's shapes, where the k splits, orientation batches,
tile counts and launch grids are those of the measurements):

  C1  Al, 99 x 99 px, 180 omega x 103 x 180 eta x 5 rings, K = 29713 (uniform 3-degree grid), sigma 2 deg:
      the FP32 CSR and the BF16 hi/lo residual matrix-core products against the FP64 oracle on a sample (forward
      rows of 4 angles of both sweep classes for coefficients on 1024 orientations, the adjoint of the counts on
      those rows over the same orientations: 2e-5 of the largest entry, 1e-5 relative L2); the transpose identity
      <A x, y> = <x, A^T y> at full size; the SPDHG iterate after 300 steps of the residual path against the FP32
      path (1e-4; the FP32 path is the one pinned to the oracle's iterates).
  C3  three slices of the multi-slice workload (128 x 128 px, 180 omega x 131 x 180 eta x 5 rings, their own
      jittered-grain grids and disc masks) batched in one run give each slice the iterates and objectives of its
      single-slice run (bf16_residual: 1e-5 / 1e-6, as the small batched test).
  C4  one slice of the differentiation workload (128 x 128 px, K = 5000, BF16 residual): the autograd Function's
      dot-product test, and the gradients of Poisson NLL(f_a A c; b) + 1e-3 TV(c) with respect to c and f_a
      against central differences of the same loss through the HIP forward (float64 loss, 16 directions)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SAMPLE_ROWS = (0, 45, 90, 135)
SAMPLE_K0, SAMPLE_KN = 12288, 1024


def rel_l2(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def rel_max(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


# ------------------------------------------------------------------------------------------------------------ C1
@pytest.fixture(scope="module")
def c1():
    from diffractom_amd import SinglePhaseForwardOperator
    from diffractom_amd.workloads import AL_RINGS, CONFIGS, make_grid, poisson_counts, voronoi_truth
    from oracle.model import OracleOperator
    print("case c1: set-up (FP32 CSR and BF16 residual operators, K = 29713)", flush=True)
    p = CONFIGS["al_c1"]
    grid, mat, cfg = make_grid("al_c1")
    ops = {q: SinglePhaseForwardOperator(cfg, mat, grid, max_gb=1024.0, normalized=True,
                                         pf_mode="sparse" if q == "fp32" else "dense", pf_precision=q)
           for q in ("fp32", "bf16_residual")}
    op = ops["fp32"]
    xt = voronoi_truth(op.K, p["n"], p["grains"], p["radius"], p["seed"], p["coeff"])
    b, _ = poisson_counts(op.direct(xt), p["mean_counts"], p["seed"])
    ks = np.arange(SAMPLE_K0, SAMPLE_K0 + SAMPLE_KN)
    rows = np.array(SAMPLE_ROWS)
    orc = OracleOperator(cfg, (4.05, 4.05, 4.05, 90, 90, 90), "cubic", AL_RINGS, 0.2, op.grid_rot_cpu[ks],
                         float(op.sigma_cpu[0]), orthonormalize=False)
    xb = np.random.default_rng(11).random((len(ks), op.Ny, op.Nx)) * op.support_mask()[None]
    fwd = orc.direct_rows(xb, rows)
    orc.set_omega_subsets([rows])
    yb = np.zeros(op.data_shape)
    yb[rows] = np.asarray(b, np.float64)[rows]
    adj = orc.adjoint(yb, subset=0)
    return {"ops": ops, "b": b, "p": p, "ks": ks, "rows": rows, "xb": xb, "fwd": fwd, "yb": yb, "adj": adj}


@pytest.mark.parametrize("precision", ["fp32", "bf16_residual"])
def test_c1_against_oracle_sample(c1, precision):
    print(f"case c1 oracle sample {precision}", flush=True)
    op = c1["ops"][precision]
    ks, rows = c1["ks"], c1["rows"]
    x = np.zeros(op.coeff_shape, np.float32)
    x[ks] = c1["xb"]
    fwd = op.direct(x)[rows]
    adj = op.adjoint(c1["yb"].astype(np.float32))[ks]
    ef, lf = rel_max(fwd, c1["fwd"]), rel_l2(fwd, c1["fwd"])
    ea, la = rel_max(adj, c1["adj"]), rel_l2(adj, c1["adj"])
    print(f"c1 {precision}: forward max {ef:.2e} L2 {lf:.2e}; adjoint max {ea:.2e} L2 {la:.2e}", flush=True)
    assert ef < 2e-5 and lf < 1e-5 and ea < 2e-5 and la < 1e-5


def test_c1_transpose_identity(c1):
    import torch
    print("case c1 transpose identity", flush=True)
    op = c1["ops"]["bf16_residual"]
    g = torch.Generator(device="cuda").manual_seed(4)
    sup = torch.from_numpy(op.support_mask()).cuda()
    x = torch.rand(op.coeff_shape, generator=g, device="cuda") * sup
    y = torch.rand(op.data_shape, generator=g, device="cuda")
    lhs = float((op.direct(x).double() * y.double()).sum())
    rhs = float((x.double() * op.adjoint(y).double()).sum())
    print(f"c1: <Ax, y> {lhs:.10e}, <x, A^T y> {rhs:.10e}", flush=True)
    assert abs(lhs / rhs - 1) < 1e-5


def test_c1_spdhg_300_steps(c1):
    from diffractom_amd import SPDHGPoisson
    print("case c1 spdhg 300 steps", flush=True)
    xs, fs = {}, {}
    for q in ("fp32", "bf16_residual"):
        op = c1["ops"][q]
        sp = SPDHGPoisson(op, subsets=c1["p"]["subsets"], seed=0)
        xs[q] = sp.run(np.zeros(op.coeff_shape, np.float32), c1["b"], 30, diagnostics_interval=10)
        fs[q] = np.array([s["f"] for s in sp.iter_stats if s["f"] is not None])
    ex = rel_l2(xs["bf16_residual"], xs["fp32"])
    ef = float(np.max(np.abs(fs["bf16_residual"] / fs["fp32"] - 1)))
    print(f"c1: x after 300 steps {ex:.2e}, objectives {ef:.2e}", flush=True)
    assert ex < 1e-4 and ef < 1e-5


# ------------------------------------------------------------------------------------------------------------ C3
def test_c3_batch_equals_single_slices_at_size():
    from diffractom_amd import SinglePhaseForwardOperator, SPDHGPoisson
    from diffractom_amd.batched import BatchedSPDHGPoisson
    from diffractom_amd.workloads import C3, c3_slices, poisson_counts
    print("case c3: 3 slices at size", flush=True)
    cfg, mat, slices = c3_slices(3)
    ops, bs, masks = [], [], []
    for i, (grid, mask, truth) in enumerate(slices):
        op = SinglePhaseForwardOperator(cfg, mat, grid, max_gb=1024.0, normalized=True, pf_mode="dense",
                                        pf_precision="bf16_residual")
        b, _ = poisson_counts(op.direct(truth), C3["mean_counts"], C3["seed"] + i)
        ops.append(op)
        bs.append(b)
        masks.append(mask)
    bat = BatchedSPDHGPoisson(ops, subsets=C3["subsets"], support=masks, seed=0)
    xs = bat.run([np.zeros(op.coeff_shape, np.float32) for op in ops], bs, 10, diagnostics_interval=10)
    f_b = bat.iter_stats[-1]["f"]
    for i, op in enumerate(ops):
        sol = SPDHGPoisson(op, subsets=C3["subsets"], support=masks[i], seed=0)
        x1 = sol.run(np.zeros(op.coeff_shape, np.float32), bs[i], 10, diagnostics_interval=10)
        err = rel_l2(xs[i], x1)
        ef = abs(f_b[i] / sol.iter_stats[-1]["f"] - 1)
        print(f"c3 slice {i} (K={op.K}, {int(masks[i].sum())} px): x {err:.2e}, f {ef:.2e}", flush=True)
        assert err < 1e-5 and ef < 1e-6


# ------------------------------------------------------------------------------------------------------------ C4
def test_c4_gradients_at_size():
    import torch
    from diffractom_amd import SinglePhaseForwardOperator
    from diffractom_amd.differentiable import IntensityLayer, forward_project, poisson_nll, total_variation
    from diffractom_amd.grid import MatrixGrid
    from diffractom_amd.workloads import al_material, jittered_grid, poisson_counts, scan_cfg, voronoi_truth
    print("case c4: one slice at size", flush=True)
    cfg = scan_cfg(128, 131, 180, 180)
    rng = np.random.default_rng(4)
    R = jittered_grid(5000, 100, 1.0, "cubic", int(rng.integers(1 << 31)))
    op = SinglePhaseForwardOperator(cfg, al_material(), MatrixGrid(R, np.deg2rad(2.0)), max_gb=1024.0,
                                    normalized=True, pf_mode="dense", pf_precision="bf16_residual")
    xt = voronoi_truth(op.K, 128, 60, 55.0, int(rng.integers(1 << 31)), coeff=(0.5, 1.5),
                       orientations=rng.choice(100, 60, replace=False))
    b = torch.from_numpy(poisson_counts(op.direct(xt), 1000.0, 0)[0]).cuda()
    supp = torch.from_numpy(op.support_mask()).cuda()
    # dot-product test of the autograd Function at size: <A x, y> through backward = <x, A^T y>
    g = torch.Generator(device="cuda").manual_seed(9)
    x = (torch.rand(op.coeff_shape, generator=g, device="cuda") * supp).requires_grad_(True)
    y = torch.rand(op.data_shape, generator=g, device="cuda")
    lhs = (forward_project(x, op).double() * y.double()).sum()
    lhs.backward()
    rhs = (x.detach().double() * x.grad.double()).sum()
    rhs_direct = (x.detach().double() * op.adjoint(y).double()).sum()
    assert torch.equal(x.grad, op.adjoint(y))
    assert abs(float(lhs) / float(rhs) - 1) < 1e-5 and float(rhs) == float(rhs_direct)
    # gradients of the C4 loss against central differences through the HIP forward
    lay = IntensityLayer(op.data_shape, op.N_peaks)
    c0 = ((0.2 + torch.rand(op.coeff_shape, generator=g, device="cuda")) * supp).float()
    f0 = torch.from_numpy(rng.uniform(0.9, 1.1, op.N_Omega)).cuda()

    def loss(c, f, grad=False):
        prm = lay.params0(device="cuda")
        prm["angle_scale"] = f
        return poisson_nll(lay.forward(forward_project(c, op), prm), b) + 1e-3 * total_variation(c)

    c = c0.clone().requires_grad_(True)
    f = f0.clone().requires_grad_(True)
    loss(c, f).backward()
    gc, gf = c.grad.double(), f.grad.double()
    h = 1e-3
    ad, fd = [], []
    with torch.no_grad():
        for i in range(16):
            v = torch.randn(op.coeff_shape, generator=g, device="cuda") * supp
            w = torch.randn(op.N_Omega, generator=g, device="cuda", dtype=torch.float64)
            lp = float(loss((c0.double() + h * v.double()).float(), f0 + h * w))
            lm = float(loss((c0.double() - h * v.double()).float(), f0 - h * w))
            fd.append((lp - lm) / (2 * h))
            ad.append(float((gc * v.double()).sum() + (gf * w).sum()))
    ad, fd = np.array(ad), np.array(fd)
    err = np.linalg.norm(ad - fd) / np.linalg.norm(fd)
    print(f"c4: dot product {float(lhs) / float(rhs) - 1:.1e}, directional derivatives vs central differences "
          f"{err:.2e}", flush=True)
    assert err < 1e-3
