"""The oracle's PDHG against an independent float64 solver: the reference's own bar (tests/test_poisson.py:189-207:
Chambolle-Pock with the best of three step balances reaches the L-BFGS-B minimiser within 1e-3 after 4000
iterations), for the plain deviance and with the l1 term lam sum(x) (smooth on x >= 0, so L-BFGS-B minimises it
directly); a run continued from its state equals one run."""
import numpy as np
from scipy.optimize import minimize

from oracle import model as O
from pairs import AL3, cfg, orientations


def _op(K=3):
    c = cfg(nx=8, ny=8, my=9, n_omega=8, angle_range=(0, 180), cor=0.0, n_eta=12)
    return O.OracleOperator(c, (4.05, 4.05, 4.05, 90, 90, 90), "cubic", AL3, 0.35, orientations(K, 5),
                            np.deg2rad(6.0), orthonormalize=False)


def _setup(op, seed=7, gain=0.05):
    rng = np.random.default_rng(seed)
    supp = op.support_mask()
    xt = rng.random(op.coeff_shape) * (rng.random(op.coeff_shape) < 0.4) * supp[None]
    ax = op.direct(xt).clip(0)
    b = rng.poisson(ax / ax.mean() * 50 / gain) * gain
    n = int(np.prod(op.coeff_shape))
    A = np.zeros((int(np.prod(op.data_shape)), n))
    e = np.zeros(n)
    for j in range(n):
        e[j] = 1.0
        A[:, j] = op.direct(e.reshape(op.coeff_shape)).ravel()
        e[j] = 0.0
    return A, b, supp


def _lbfgs(A, b, supp, shape, gain, lam):
    eps = float(np.float32(1e-3))
    bb = b.ravel()

    def fg(x):
        mu = np.maximum(A @ x, 0) + eps * gain
        v = mu - bb + np.where(bb > 0, bb * np.log(np.where(bb > 0, bb, 1) / mu), 0)
        return float(v.sum() / gain + lam * x.sum()), A.T @ ((1 - bb / mu) / gain) + lam

    free = np.broadcast_to(supp, shape).ravel() & (A.sum(axis=0) > 0)
    r = minimize(fg, np.full(A.shape[1], 0.1) * free, jac=True, method="L-BFGS-B",
                 bounds=[(0, None) if f else (0, 0) for f in free],
                 options={"maxiter": 20000, "ftol": 1e-15, "gtol": 1e-12})
    return r.fun


def test_pdhg_reaches_the_minimiser():
    op = _op()
    gain = 0.05
    A, b, supp = _setup(op, gain=gain)
    for lam, kind in ((0.0, "nonneg"), (40.0, "nonneg_l1")):
        f_ref = _lbfgs(A, b, supp, op.coeff_shape, gain, lam)
        best = None
        for rho in (1e-4, 1e-2, 1.0):
            pd = O.OraclePDHG(op, prox_kind=kind, lam=lam, rho=rho)
            x = pd.run(np.zeros(op.coeff_shape), b, 4000, gain=gain)
            val = pd.final_f + lam * x.sum()
            best = val if best is None else min(best, val)
        assert best <= f_ref * (1 + 1e-3)


def test_pdhg_resumes_from_its_state():
    op = _op()
    A, b, supp = _setup(op, seed=3, gain=0.05)
    one = O.OraclePDHG(op, rho=1e-2)
    x1 = one.run(np.zeros(op.coeff_shape), b, 30, gain=0.05)
    two = O.OraclePDHG(op, rho=1e-2)
    x2 = two.run(np.zeros(op.coeff_shape), b, 18, gain=0.05)
    x2 = two.run(x2, b, 12, gain=0.05, state=two.state)
    np.testing.assert_allclose(x2, x1, rtol=1e-12, atol=1e-12 * np.abs(x1).max())
    assert two.state["it"] == 30
