"""Many slices in one run (BatchedSPDHGPoisson): slices with different orientation grids (K) and support masks,
every kernel of a step launched once for all of them, give each slice the iterates and objectives of SPDHGPoisson run
on that slice alone with the same seed (bf16_residual: relative L2 1e-5, objectives 1e-6; plain bf16: 1e-3, 1e-4;
the single-slice path is pinned to the FP64 oracle in test_gpu_mfma.py)."""
import numpy as np
import pytest

from pairs import build, cfg, rel_l2

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", ["bf16_residual", "bf16"])
def test_batch_equals_single_slices(precision):
    import torch
    from diffractom_amd import SPDHGPoisson
    from diffractom_amd.batched import BatchedSPDHGPoisson
    print(f"case batched {precision}", flush=True)
    c = cfg(nx=20, ny=20, my=21, n_omega=24, angle_range=(0, 180), cor=0.3)
    ops, bs, sups = [], [], []
    rng = np.random.default_rng(11)
    yy, xx = np.mgrid[:20, :20]
    for i, (K, r) in enumerate([(20, 9.0), (75, 6.0), (130, None)]):
        op, orc, _ = build(c, K=K, seed=3 + i, pf_mode="dense", pf_precision=precision)
        sup = "fov" if r is None else ((xx - 9.5) ** 2 + (yy - 9.5) ** 2 <= r * r).astype(np.uint8)
        m = op.support_mask() if r is None else sup
        xt = (rng.random(op.coeff_shape) * (rng.random(op.coeff_shape) < 0.3) * m[None]).astype(np.float32)
        ax = orc.direct(xt)
        bs.append(rng.poisson(ax / max(ax.mean(), 1e-9) * 40).astype(np.float32))
        ops.append(op)
        sups.append(sup)
    bat = BatchedSPDHGPoisson(ops, subsets=4, support=sups, seed=5)
    xs = bat.run([np.zeros(op.coeff_shape, np.float32) for op in ops], bs, 5, diagnostics_interval=5)
    assert bat.Kc == sum(op.K_stride for op in ops)
    f_b = bat.iter_stats[-1]["f"]
    for i, op in enumerate(ops):
        sol = SPDHGPoisson(op, subsets=4, support=sups[i], seed=5)
        x1 = sol.run(np.zeros(op.coeff_shape, np.float32), bs[i], 5, diagnostics_interval=5)
        err = rel_l2(xs[i], x1)
        print(f"  slice {i} (K={op.K}): rel L2 {err:.2e}, f {f_b[i]:.6f} vs {sol.iter_stats[-1]['f']:.6f}", flush=True)
        # the single-slice run splits K over partial sums where the batch writes directly: the same products summed
        # in another order; plain bf16's iterates amplify that more (its error band against the oracle is 3e-3)
        tol_x, tol_f = (1e-5, 1e-6) if precision == "bf16_residual" else (1e-3, 1e-4)
        assert err < tol_x
        assert abs(f_b[i] / sol.iter_stats[-1]["f"] - 1) < tol_f


def test_batch_autotune_keeps_the_iterates():
    """The batch's projector autotune (times every forward / backward candidate on the batched buffers, the
    backprojection on copies of x and z) leaves the solver state untouched, and the tuned steps give the same iterates
    as the default kernels up to float32 summation order."""
    import torch
    from diffractom_amd.batched import BatchedSPDHGPoisson
    print("case batched autotune", flush=True)
    c = cfg(nx=20, ny=20, my=21, n_omega=24, angle_range=(0, 180), cor=0.3)
    ops, bs = [], []
    rng = np.random.default_rng(5)
    for i, K in enumerate((40, 90)):
        op, orc, _ = build(c, K=K, seed=7 + i, pf_mode="dense", pf_precision="bf16_residual")
        xt = (rng.random(op.coeff_shape) * op.support_mask()[None]).astype(np.float32)
        ax = orc.direct(xt)
        bs.append(rng.poisson(ax / max(ax.mean(), 1e-9) * 40).astype(np.float32))
        ops.append(op)
    ref = BatchedSPDHGPoisson(ops, subsets=4, seed=2)
    ref.start(None, bs)
    ref.advance(10)
    x_ref = ref.xk.clone()
    bat = BatchedSPDHGPoisson(ops, subsets=4, seed=2)
    bat.start(None, bs)
    x0, z0 = bat.xk.clone(), bat.zk.clone()
    tuning = bat.autotune(reps=1)
    assert "radon_forward" in tuning and "radon_backward_spdhg" in tuning
    assert torch.equal(bat.xk, x0) and torch.equal(bat.zk, z0)
    bat.advance(10)
    assert rel_l2(bat.xk.cpu().numpy(), x_ref.cpu().numpy()) < 1e-5


@pytest.mark.parametrize("my", [21, 131])
@pytest.mark.parametrize("precision", ["bf16_residual", "fp16"])
def test_batch_of_block_sparse_slices(precision, my):
    """The grouped block-sparse PF products (dfm_pfblk_grouped: one launch for every slice's blocks) give each
    slice the iterates and objectives of its single-slice run on the same blocks (1e-5 / 1e-6); at 131 detector rows
    (two groups of 5 y tiles) the grouped launch runs its form with the tile count compiled in."""
    from diffractom_amd import SPDHGPoisson
    from diffractom_amd.batched import BatchedSPDHGPoisson
    print(f"case batched blocks {precision} my={my}", flush=True)
    c = cfg(nx=20, ny=20, my=my, n_omega=24, angle_range=(0, 180), cor=0.3)
    ops, bs = [], []
    rng = np.random.default_rng(12)
    for i, K in enumerate((20, 75, 130)):
        op, orc, _ = build(c, K=K, seed=3 + i, pf_mode="sparse", pf_precision=precision)
        assert op.blk_batches is not None
        xt = (rng.random(op.coeff_shape) * (rng.random(op.coeff_shape) < 0.3) * op.support_mask()[None]
              ).astype(np.float32)
        ax = orc.direct(xt)
        bs.append(rng.poisson(ax / max(ax.mean(), 1e-9) * 40).astype(np.float32))
        ops.append(op)
    bat = BatchedSPDHGPoisson(ops, subsets=4, seed=5)
    assert bat.blocks
    xs = bat.run([np.zeros(op.coeff_shape, np.float32) for op in ops], bs, 5, diagnostics_interval=5)
    f_b = bat.iter_stats[-1]["f"]
    for i, op in enumerate(ops):
        sol = SPDHGPoisson(op, subsets=4, seed=5)
        x1 = sol.run(np.zeros(op.coeff_shape, np.float32), bs[i], 5, diagnostics_interval=5)
        err = rel_l2(xs[i], x1)
        print(f"  slice {i} (K={op.K}): rel L2 {err:.2e}", flush=True)
        assert err < 1e-5
        assert abs(f_b[i] / sol.iter_stats[-1]["f"] - 1) < 1e-6


def test_batched_graph_replay_matches_eager():
    """Batched steps replayed from HIP graphs (every ordered subset pair captured up front, then the epoch order
    replayed with the solver's own eager work in between) give the eager iterates bit for bit, over several epochs."""
    import torch
    from diffractom_amd.batched import BatchedSPDHGPoisson
    print("case batched graphs", flush=True)
    c = cfg(nx=20, ny=20, my=21, n_omega=24, angle_range=(0, 180), cor=0.3)
    ops, bs = [], []
    rng = np.random.default_rng(21)
    for i, K in enumerate((40, 90)):
        op, orc, _ = build(c, K=K, seed=9 + i, pf_mode="sparse", pf_precision="bf16_residual")
        xt = (rng.random(op.coeff_shape) * op.support_mask()[None]).astype(np.float32)
        ax = orc.direct(xt)
        bs.append(rng.poisson(ax / max(ax.mean(), 1e-9) * 40).astype(np.float32))
        ops.append(op)
    res = []
    for graphs in (False, True):
        bat = BatchedSPDHGPoisson(ops, subsets=4, seed=3, graphs=graphs)
        bat.start(None, bs)
        if graphs:
            assert bat.capture_all() == 12
        for _ in range(5):
            bat.advance(4)
            torch.cuda.synchronize()
        res.append(bat.xk.clone())
    assert torch.equal(res[0], res[1])
