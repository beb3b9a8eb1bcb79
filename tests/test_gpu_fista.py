"""FISTAPoisson on the GPU vs the FP64 oracle restatement of fista_poisson.py (same start phase, metric, descent test,
restarts and refreshes): the objective of every iteration within 1e-4 and the iterate within 1e-3 (relative L2)
after 40 iterations (20 weighted-least-squares, 20 Poisson), the backtracking counts equal; nonneg_l1 with weights
and background; the uniform start; NumPy and torch x0; argument errors."""
import numpy as np
import pytest

from oracle.model import OracleFISTA
from pairs import build, cfg, rel_l2

pytestmark = pytest.mark.gpu


def _problem(K=12, seed=7, gain=0.05):
    op, orc, _ = build(cfg(nx=16, ny=16, my=17, n_omega=24, angle_range=(0, 180), cor=0.0), K=K, sigma_deg=4.0,
                       pf_mode="sparse")
    rng = np.random.default_rng(seed)
    xt = (rng.random(op.coeff_shape) * (rng.random(op.coeff_shape) < 0.2) * op.support_mask()[None]
          ).astype(np.float32)
    ax = orc.direct(xt)
    b = (rng.poisson(ax / ax.mean() * 50 / gain) * gain).astype(np.float32)
    return op, orc, b


@pytest.mark.parametrize("kind,extras,start", [("nonneg", False, "wls"), ("nonneg_l1", True, "wls"),
                                               ("nonneg", False, "uniform")])
def test_fista_matches_oracle(kind, extras, start):
    from diffractom_amd import FISTAPoisson
    print(f"case kind={kind} extras={extras} start={start}", flush=True)
    op, orc, b = _problem()
    kw = {"gain": 0.05}
    if extras:
        rng = np.random.default_rng(3)
        kw["weights"] = (rng.random(op.N_seg) > 0.2).astype(np.float32)
        kw["background"] = np.full(op.N_seg, 0.3, np.float32)
    lam = 20.0 if kind == "nonneg_l1" else 0.0
    fi = FISTAPoisson(op, prox_kind=kind, lam=lam, start=start)
    x = fi.run(np.zeros(op.coeff_shape, np.float32), b, 40, **kw)
    ref = OracleFISTA(orc, prox_kind=kind, lam=lam, start=start)
    x_ref = ref.run(np.zeros(op.coeff_shape), b, 40, **kw)
    obj = [s["obj"] for s in fi.iter_stats]
    obj_ref = [s["obj"] for s in ref.iter_stats]
    np.testing.assert_allclose(obj, obj_ref, rtol=1e-4)
    assert [s["backtracks"] for s in fi.iter_stats] == [s["backtracks"] for s in ref.iter_stats]
    assert rel_l2(x, x_ref) < 1e-3
    assert abs(fi.final_stats["final_obj"] / ref.final_obj - 1) < 1e-4
    assert abs(fi.final_stats["adjoint_scale"] - 1) < 1e-4
    assert x.min() >= 0 and (x * ~op.support_mask()[None]).max() == 0


def test_fista_x0_kinds_and_errors():
    import torch
    from diffractom_amd import FISTAPoisson
    op, _, b = _problem(seed=4)
    x0 = np.zeros(op.coeff_shape, np.float32)
    one = FISTAPoisson(op)
    x_one = one.run(x0, b, 25, gain=0.05)
    assert x_one is x0 and one.state["it"] == 25 and one.final_stats["niter"] == 25
    xt = torch.zeros(op.coeff_shape, dtype=torch.float32, device="cuda")
    out = FISTAPoisson(op).run(xt, b, 25, gain=0.05)
    assert out is xt
    np.testing.assert_allclose(xt.cpu().numpy(), x_one, rtol=1e-6, atol=1e-6 * np.abs(x_one).max())
    two = FISTAPoisson(op)
    two.run(x_one.copy(), b, 10, gain=0.05, state=one.state)
    assert two.state["it"] == 35 and two.iter_stats[0]["iter"] == 26
    with pytest.raises(ValueError):
        FISTAPoisson(op, prox_kind="tv")
    with pytest.raises(ValueError):
        FISTAPoisson(op, start="random")
    with pytest.raises(ValueError):
        FISTAPoisson(op).run(x0, b, 2, gain=np.zeros(op.N_seg, np.float32))
