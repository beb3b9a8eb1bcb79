"""The C0 workload at full size (99 x 99 pixels, 180 omega x 103 translations x 180 eta x 5 rings, K = 2000,
sigma 2 deg) on the HIP path against the FP64 oracle, for the FP32 CSR products and the headline BF16 matrix-core
products with the hi/lo residual correction:

  * the forward projection of the Voronoi truth and the adjoint of the Poisson counts, every entry, against the
    oracle's evaluated here (the counts are drawn from the oracle's float64 forward, as the fixture script did);
  * 200 SPDHG subset steps (10 subsets, seed 0) against the committed oracle run (golden/c0_oracle.npz, written by
    golden/make_c0_fixture.py): the objectives after 5, 10, 15 and 20 epochs, the iterate's seeded entry sample,
    norms, per-orientation and per-pixel sums.

Tolerances: the forward / adjoint within 2e-5 of the largest entry and 1e-5 relative L2 (the reference's GPU-vs-FP64
bar, tests/test_differentiable.py:36-41); objectives within 1e-5; the iterate within 1e-4 (relative L2 of the sample
and of the sums)."""
import numpy as np
import pytest

import c0

pytestmark = pytest.mark.gpu

PRECISIONS = ["fp32", "bf16_residual"]


def rel_l2(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def rel_max(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


@pytest.fixture(scope="module")
def prob():
    print("case c0: FP64 oracle operator, forward and adjoint (about a minute on the host cores)", flush=True)
    cfg, U, sigma, xt, p, (grid, mat) = c0.problem()
    fx = dict(np.load(c0.FIXTURE, allow_pickle=False))
    orc = c0.oracle_operator(cfg, U, sigma)
    ax = orc.direct(xt)
    b = c0.counts(ax, p)
    # the oracle here reproduces the committed oracle figures (the counts exactly)
    assert float(np.asarray(b, np.float64).sum()) == float(fx["b_sum"])
    assert rel_l2(ax.ravel()[fx["fwd_idx"]], fx["fwd_val"]) < 1e-12
    adj = orc.adjoint(b)
    assert rel_l2(adj.ravel()[fx["adj_idx"]], fx["adj_val"]) < 1e-12
    print("case c0: oracle done", flush=True)
    return {"cfg": cfg, "grid": grid, "mat": mat, "xt": xt, "p": p, "ax": ax, "b": b, "adj": adj, "fx": fx}


@pytest.fixture(scope="module")
def ops(prob):
    from diffractom_amd import SinglePhaseForwardOperator
    made = {}

    def get(precision):
        if precision not in made:
            for q in list(made):                     # one operator at a time
                made.pop(q).free_memory()
            made[precision] = SinglePhaseForwardOperator(
                prob["cfg"], prob["mat"], prob["grid"], max_gb=64.0, normalized=True,
                pf_mode="sparse" if precision == "fp32" else "dense", pf_precision=precision)
        return made[precision]
    return get


@pytest.mark.parametrize("precision", PRECISIONS)
def test_c0_forward_adjoint_full_size(prob, ops, precision):
    print(f"case c0 forward/adjoint {precision}", flush=True)
    op = ops(precision)
    fwd = op.direct(prob["xt"])
    adj = op.adjoint(prob["b"])
    ef, ea = rel_max(fwd, prob["ax"]), rel_max(adj, prob["adj"])
    lf, la = rel_l2(fwd, prob["ax"]), rel_l2(adj, prob["adj"])
    print(f"c0 {precision}: forward max {ef:.2e} L2 {lf:.2e}; adjoint max {ea:.2e} L2 {la:.2e}", flush=True)
    assert ef < 2e-5 and ea < 2e-5 and lf < 1e-5 and la < 1e-5


@pytest.mark.parametrize("precision", PRECISIONS)
def test_c0_spdhg_200_steps(prob, ops, precision):
    from diffractom_amd import SPDHGPoisson
    print(f"case c0 spdhg {precision}", flush=True)
    op = ops(precision)
    fx, p = prob["fx"], prob["p"]
    sp = SPDHGPoisson(op, subsets=p["subsets"], seed=0)
    x = sp.run(np.zeros(op.coeff_shape, np.float32), prob["b"], c0.EPOCHS, diagnostics_interval=c0.DIAG)
    f = np.array([s["f"] for s in sp.iter_stats if s["f"] is not None])
    assert np.array_equal(np.array([s["epoch"] for s in sp.iter_stats if s["f"] is not None]), fx["spdhg_epochs"])
    ef = float(np.max(np.abs(f - fx["spdhg_f"]) / np.abs(fx["spdhg_f"])))
    assert abs(sp.rho_used - float(fx["rho"])) <= 1e-6 * float(fx["rho"])
    s = c0.summary(x, 3)
    assert np.array_equal(s["idx"], fx["x_idx"])
    es = rel_l2(s["val"], fx["x_val"])
    e0, e1 = rel_l2(s["sum0"], fx["x_sum0"]), rel_l2(s["sum_rest"], fx["x_sum_rest"])
    en = abs(s["l2"] - float(fx["x_l2"])) / float(fx["x_l2"])
    print(f"c0 {precision}: objectives {ef:.2e}, iterate sample {es:.2e}, per-pixel sums {e0:.2e}, "
          f"per-orientation sums {e1:.2e}, norm {en:.2e}", flush=True)
    assert ef < 1e-5
    assert es < 1e-4 and e0 < 1e-4 and e1 < 1e-4 and en < 1e-5
