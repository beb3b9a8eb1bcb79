"""The large hexagonal workload (Ti, 8 rings, 360 omega x 261 translations x 360 eta, 256 x 256 pixels, 59,440
orientations, sigma 2.5 deg) with the PF generated per call on the device in fp16 blocks (pf_mode "generated",
pf_precision "fp16": no PF stored beyond the block counts; the FP32 CSR would take about 48 GB):

  * the forward rows of 2 angles for coefficients on a block of 256 orientations straddling two orientation batches,
    and the adjoint of the counts on those rows over the block, against the FP64 oracle at fp16's bound (unit
    roundoff 2^-11 on both operands of each product: 2e-3 of the largest entry, 8e-4 relative L2);
  * the transpose identity <A x, y> = <x, A^T y> over everything (1e-5: both directions use the same fp16 entries,
    so only the FP32 accumulation order separates them);
  * the subset partition: the 10 subset adjoints add up to the full adjoint (1e-6) and a subset forward writes its
    own rows only, equal to those rows of the full forward (1e-6)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def rel_max(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


@pytest.fixture(scope="module")
def c2g():
    import torch
    from diffractom_amd.workloads import c2_problem
    print("case c2 generated fp16: set-up (block counts only)", flush=True)
    op, xk, b, p = c2_problem(pf_mode="generated", pf_precision="fp16")
    assert op.gen_batches is not None
    del xk
    torch.cuda.empty_cache()
    return op, b, p


def test_c2_generated_against_oracle_sample(c2g):
    import torch
    from diffractom_amd.workloads import TI_RINGS
    from oracle.model import OracleOperator
    op, b, _ = c2g
    kb0 = op.batches[0]["K_batch"]
    ks = np.arange(kb0 - 128, kb0 + 128)
    rows = np.array([10, 100])
    print(f"case c2 generated oracle sample: orientations {ks[0]}..{ks[-1]}, angles {rows.tolist()}", flush=True)
    orc = OracleOperator(op.cfg, (2.95, 2.95, 4.68, 90, 90, 120), "hexagonal", TI_RINGS, 0.2, op.grid_rot_cpu[ks],
                         float(op.sigma_cpu[0]), orthonormalize=False)
    xb = np.random.default_rng(3).random((len(ks), op.Ny, op.Nx)) * op.support_mask()[None]
    ref_f = orc.direct_rows(xb, rows)
    orc.set_omega_subsets([rows])
    yb = np.zeros(op.data_shape)
    rt = torch.from_numpy(rows).cuda()
    yb[rows] = b[rt].double().cpu().numpy()
    ref_a = orc.adjoint(yb, subset=0)
    x = torch.zeros(op.coeff_shape, dtype=torch.float32, device="cuda")
    x[int(ks[0]):int(ks[-1]) + 1] = torch.from_numpy(xb.astype(np.float32)).cuda()
    fwd = op.direct(x)[rt].double().cpu().numpy()
    del x
    y = torch.zeros(op.data_shape, dtype=torch.float32, device="cuda")
    y[rt] = b[rt]
    adj = op.adjoint(y)[int(ks[0]):int(ks[-1]) + 1].double().cpu().numpy()
    del y
    ef, lf, ea, la = rel_max(fwd, ref_f), rel_l2(fwd, ref_f), rel_max(adj, ref_a), rel_l2(adj, ref_a)
    print(f"c2 generated: forward max {ef:.2e} L2 {lf:.2e}; adjoint max {ea:.2e} L2 {la:.2e}", flush=True)
    assert ef < 2e-3 and lf < 8e-4 and ea < 2e-3 and la < 8e-4


def test_c2_generated_transpose_and_subsets(c2g):
    import torch
    op, b, p = c2g
    print("case c2 generated: transpose identity and subset partition", flush=True)
    g = torch.Generator(device="cuda").manual_seed(6)
    sup = torch.from_numpy(op.support_mask()).cuda()
    x = torch.rand(op.coeff_shape, generator=g, device="cuda") * sup
    y = torch.rand(op.data_shape, generator=g, device="cuda")
    ax = op.direct(x)
    lhs = float((ax.double() * y.double()).sum())
    aty = op.adjoint(y)
    rhs = float((x.double() * aty.double()).sum())
    print(f"c2 generated: <Ax, y> {lhs:.10e}, <x, A^T y> {rhs:.10e}", flush=True)
    assert abs(lhs / rhs - 1) < 1e-5
    subs = op.set_omega_subsets(p["subsets"])
    acc = torch.zeros_like(aty, dtype=torch.float64)
    part = torch.empty_like(aty)
    for m in range(len(subs)):
        op.adjoint_cl(y, part, subset=m)
        acc += part.double()
    e = float(torch.linalg.vector_norm(acc - aty.double()) / torch.linalg.vector_norm(aty.double()))
    print(f"c2 generated: sum of subset adjoints vs full adjoint {e:.2e}", flush=True)
    assert e < 1e-6
    d = torch.full(op.data_shape, -7.0, device="cuda")
    op.direct_cl(x, d, subset=3)
    rows = torch.from_numpy(subs[3]).cuda()
    mask = torch.ones(op.N_Omega, dtype=torch.bool, device="cuda")
    mask[rows] = False
    assert bool((d[mask] == -7.0).all())
    assert rel_l2(d[rows].double().cpu().numpy(), ax[rows].double().cpu().numpy()) < 1e-6
