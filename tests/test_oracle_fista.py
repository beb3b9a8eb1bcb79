"""The oracle's FISTA (SQS metric, backtracking) against an independent float64 solver: the L-BFGS-B minimiser of the
deviance (and of the deviance + lam sum(x), smooth on x >= 0) reached within 1e-3 (the reference's bar for its
Poisson solvers, tests/test_poisson.py:189-207); the objective after the start phase never increases by more than
the descent test allows between restarts; a run continued from its state continues the descent."""
import numpy as np

from oracle import model as O
from test_oracle_pdhg import _lbfgs, _op, _setup


def test_fista_reaches_the_minimiser():
    op = _op()
    gain = 0.05
    A, b, supp = _setup(op, gain=gain)
    for lam, kind in ((0.0, "nonneg"), (40.0, "nonneg_l1")):
        f_ref = _lbfgs(A, b, supp, op.coeff_shape, gain, lam)
        fi = O.OracleFISTA(op, prox_kind=kind, lam=lam)
        fi.run(np.zeros(op.coeff_shape), b, 400, gain=gain)
        assert fi.final_obj <= f_ref * (1 + 1e-3)
        obj = np.array([s["obj"] for s in fi.iter_stats if s["phase"] == "poisson"])
        assert obj[-1] <= obj[0]


def test_fista_resumes_from_its_state():
    """A continued run starts from the state's momentum point with the metric taken at the resume point (as the
    reference recomputes it there), so it is not bit-identical to one run: it continues the count and the descent
    and ends as low as one run (to 1e-3)."""
    op = _op()
    _, b, _ = _setup(op, seed=3, gain=0.05)
    one = O.OracleFISTA(op)
    one.run(np.zeros(op.coeff_shape), b, 120, gain=0.05)
    two = O.OracleFISTA(op)
    x2 = two.run(np.zeros(op.coeff_shape), b, 60, gain=0.05)
    f_mid = two.final_obj
    two.run(x2, b, 60, gain=0.05, state=two.state)
    assert two.state["it"] == 120 and two.iter_stats[0]["iter"] == 61
    assert two.final_obj <= f_mid
    assert abs(two.final_obj / one.final_obj - 1) < 1e-3
