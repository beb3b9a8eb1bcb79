"""SPDHGPoisson on the GPU vs the FP64 oracle restatement of spdhg_poisson.py (same subset order from the same
seed): iterate, objective, dual state; warm start; the same result for a NumPy and a torch x0.
Tolerance: FP32 iterates within 1e-4 (relative L2) of the FP64 iterates after 8 epochs, objectives within 1e-5."""
import numpy as np
import pytest

from oracle.model import OracleSPDHG
from pairs import build, cfg, rel_l2

pytestmark = pytest.mark.gpu


def _problem(K=12, seed=7, gain=None):
    op, orc, _ = build(cfg(nx=16, ny=16, my=17, n_omega=24, angle_range=(0, 180), cor=0.0), K=K, sigma_deg=4.0,
                       pf_mode="sparse")
    rng = np.random.default_rng(seed)
    xt = (rng.random(op.coeff_shape) * (rng.random(op.coeff_shape) < 0.2) * op.support_mask()[None]
          ).astype(np.float32)
    ax = orc.direct(xt)
    b = rng.poisson(ax / ax.mean() * 50).astype(np.float32)
    return op, orc, b


@pytest.mark.parametrize("subsets,weights", [(4, None), (3, "eta")])
def test_spdhg_matches_oracle(subsets, weights):
    from diffractom_amd import SPDHGPoisson
    print(f"case subsets={subsets} weights={weights}", flush=True)
    op, orc, b = _problem()
    w = None
    if weights == "eta":
        w = np.ones((op.N_eta, op.N_peaks), np.float32)
        w[5:9] = 0
    sp = SPDHGPoisson(op, subsets=subsets, seed=1)
    x = sp.run(np.zeros(op.coeff_shape, np.float32), b, 8, weights=w, diagnostics_interval=2)
    ref = OracleSPDHG(orc, subsets=subsets, seed=1)
    x_ref = ref.run(np.zeros(op.coeff_shape), b, 8, weights=w, diagnostics_interval=2)
    assert abs(sp.rho_used / ref.rho_used - 1) < 1e-6
    assert rel_l2(x, x_ref) < 1e-4
    f = [s["f"] for s in sp.iter_stats if s["f"] is not None]
    f_ref = [s["f"] for s in ref.iter_stats if s["f"] is not None]
    np.testing.assert_allclose(f, f_ref, rtol=1e-5)
    assert rel_l2(sp.state["y"], ref.state["y"]) < 1e-4
    assert rel_l2(sp.state["z"], ref.state["z"]) < 1e-4


def test_spdhg_resume_and_torch_x0():
    """tests/test_subsets.py:148-171: two runs continued from solver.state equal one run; a torch x0 gives the
    NumPy x0's iterates."""
    import torch
    from diffractom_amd import SPDHGPoisson
    op, _, b = _problem()
    one = SPDHGPoisson(op, subsets=3, seed=1)
    x_one = one.run(np.zeros(op.coeff_shape, np.float32), b, 6, diagnostics_interval=2)
    xt = torch.zeros(op.coeff_shape, dtype=torch.float32, device="cuda")
    tor = SPDHGPoisson(op, subsets=3, seed=1)
    out = tor.run(xt, b, 6, diagnostics_interval=2)
    assert out is xt
    assert np.array_equal(xt.cpu().numpy(), x_one)
    two = SPDHGPoisson(op, subsets=3, seed=1)
    x = two.run(np.zeros(op.coeff_shape, np.float32), b, 4, diagnostics_interval=4)
    x = two.run(x, b, 2, diagnostics_interval=2, state=two.state)
    assert two.state["epoch"] == 6
    np.testing.assert_allclose(x, x_one, rtol=1e-5, atol=1e-6 * np.abs(x_one).max())
    st = {"y": two.state["y"], "epoch": 6}
    three = SPDHGPoisson(op, subsets=3, seed=2)
    three.run(x.copy(), b, 1, state=st)
    assert three.state["epoch"] == 7


def test_spdhg_several_batches_matches_oracle():
    """Three orientation batches (the backprojection-update and forward of each batch in turn): iterates within 1e-4
    (relative L2) of the FP64 oracle after 6 epochs, objectives within 1e-5."""
    from diffractom_amd import SPDHGPoisson
    op, orc, _ = build(cfg(nx=16, ny=16, my=17, n_omega=24, angle_range=(0, 180), cor=0.0), K=24, sigma_deg=4.0,
                       pf_mode="sparse", max_gb=8.5 * 24 * 36 * 3 * 4 / 1024 ** 3)
    assert len(op.batches) == 3
    rng = np.random.default_rng(7)
    xt = rng.random(op.coeff_shape) * (rng.random(op.coeff_shape) < 0.2) * op.support_mask()[None]
    ax = orc.direct(xt)
    b = rng.poisson(ax / ax.mean() * 50).astype(np.float32)
    sp = SPDHGPoisson(op, subsets=4, seed=3)
    x = sp.run(np.zeros(op.coeff_shape, np.float32), b, 6, diagnostics_interval=3)
    ref = OracleSPDHG(orc, subsets=4, seed=3)
    x_ref = ref.run(np.zeros(op.coeff_shape), b, 6, diagnostics_interval=3)
    assert rel_l2(x, x_ref) < 1e-4
    np.testing.assert_allclose([s["f"] for s in sp.iter_stats if s["f"] is not None],
                               [s["f"] for s in ref.iter_stats if s["f"] is not None], rtol=1e-5)


@pytest.mark.parametrize("tv,tv_prob", [(0.5, None), (2.0, 0.4)])
def test_spdhg_tv_matches_oracle(tv, tv_prob):
    """The TV block (spdhg_poisson.py:108-193): the same block sequence from the same seed, iterates within 1e-4
    (relative L2) of the FP64 oracle after 6 epochs (the oracle's TV is pinned against an independent
    Chambolle-Pock in test_oracle_tv.py), objectives (deviance + tv TV) within 1e-5, TV duals within 1e-4; a run
    continued from solver.state equals one run."""
    from diffractom_amd import SPDHGPoisson
    print(f"case tv={tv} tv_prob={tv_prob}", flush=True)
    op, orc, b = _problem()
    sp = SPDHGPoisson(op, subsets=3, seed=2, tv=tv, tv_prob=tv_prob)
    x = sp.run(np.zeros(op.coeff_shape, np.float32), b, 6, diagnostics_interval=3)
    ref = OracleSPDHG(orc, subsets=3, seed=2, tv=tv, tv_prob=tv_prob)
    x_ref = ref.run(np.zeros(op.coeff_shape), b, 6, diagnostics_interval=3)
    assert rel_l2(x, x_ref) < 1e-4
    f = [s["f"] for s in sp.iter_stats if s["f"] is not None]
    f_ref = [s["f"] for s in ref.iter_stats if s["f"] is not None]
    np.testing.assert_allclose(f, f_ref, rtol=1e-5)
    for nm in ("qx", "qy"):
        assert rel_l2(sp.state[nm], ref.state[nm]) < 1e-4
    # resume (tests/test_spdhg_tv.py): 4 + 2 epochs continued from solver.state (z, y, qx, qy) equal 6 epochs
    two = SPDHGPoisson(op, subsets=3, seed=2, tv=tv, tv_prob=tv_prob)
    x2 = two.run(np.zeros(op.coeff_shape, np.float32), b, 4, diagnostics_interval=4)
    x2 = two.run(x2, b, 2, diagnostics_interval=2, state=two.state)
    np.testing.assert_allclose(x2, x, rtol=1e-5, atol=1e-6 * np.abs(x).max())
    with pytest.raises(ValueError):
        SPDHGPoisson(op, subsets=3, tv=tv).run(x2, b, 1, state={"y": two.state["y"], "epoch": 6})


@pytest.mark.parametrize("tv", [0.0, 0.5])
def test_graph_replay_matches_eager(tv):
    """Subset steps replayed from captured HIP graphs (captured at the second occurrence of a (subset, next subset)
    pair, replayed from then on, eager first occurrences, epoch-end full forwards and objectives in between): the
    iterate after every epoch equals the eager run's bit for bit."""
    from diffractom_amd import SPDHGPoisson
    print(f"case graphs tv={tv}", flush=True)
    op, _, b = _problem(K=16)
    xs = {}
    for graphs in (False, True):
        sp = SPDHGPoisson(op, subsets=4, seed=5, tv=tv)
        sp.graphs = graphs
        seen = []
        sp.run(np.zeros(op.coeff_shape, np.float32), b, 12, diagnostics_interval=3,
               callback=lambda e, x, st: seen.append(np.array(x, copy=True)) and False)
        xs[graphs] = seen
        if graphs:
            assert len(sp._graph_cache) > 0
    assert len(xs[True]) == len(xs[False]) == 12
    for a, c in zip(xs[True], xs[False]):
        assert np.array_equal(a, c)


def test_graph_replay_across_resumed_runs():
    """A solver with graphs on runs twice (4 epochs, then 2 more continued from solver.state): the second run
    allocates its arrays anew, so it must capture its own graphs rather than replay the first run's (which hold the
    first run's buffer addresses). The result equals the same 4 + 2 epochs run eagerly, bit for bit, and so does a
    third run on inputs allocated between the runs."""
    import torch
    from diffractom_amd import SPDHGPoisson
    print("case graphs resumed", flush=True)
    op, _, b = _problem(K=16)
    out = {}
    for graphs in (False, True):
        sp = SPDHGPoisson(op, subsets=4, seed=5)
        sp.graphs = graphs
        x = sp.run(np.zeros(op.coeff_shape, np.float32), b, 4, diagnostics_interval=2)
        junk = [torch.empty(1 << 20, device="cuda") for _ in range(4)]     # the allocator hands out other blocks
        x = sp.run(x, b, 2, diagnostics_interval=2, state=sp.state)
        del junk
        x3 = sp.run(np.zeros(op.coeff_shape, np.float32), b * 2, 3, diagnostics_interval=3)
        out[graphs] = (np.array(x, copy=True), np.array(x3, copy=True), [s["f"] for s in sp.iter_stats])
        if graphs:
            assert len(sp._graph_cache) > 0
    assert np.array_equal(out[True][0], out[False][0])
    assert np.array_equal(out[True][1], out[False][1])
    assert out[True][2] == out[False][2]
