"""The oracle's SPDHG against mathematics: the dual prox's optimality condition, convergence to the minimiser of an
independent float64 solver (L-BFGS-B on the explicit matrix; the reference's own bar, tests/test_subsets.py:130-145,
is SPDHG <= a long FISTA run x (1 + 2e-3)), and the power iteration against the explicit spectrum."""
import numpy as np
import pytest
from scipy.optimize import minimize

from oracle import model as O
from pairs import AL3, cfg, orientations


def _op(K=4):
    c = cfg(nx=8, ny=8, my=9, n_omega=8, angle_range=(0, 180), cor=0.0, n_eta=12)
    return O.OracleOperator(c, (4.05, 4.05, 4.05, 90, 90, 90), "cubic", AL3, 0.35, orientations(K, 5),
                            np.deg2rad(6.0), orthonormalize=False)


def _matrix(op):
    n = int(np.prod(op.coeff_shape))
    A = np.zeros((int(np.prod(op.data_shape)), n))
    e = np.zeros(n)
    for j in range(n):
        e[j] = 1.0
        A[:, j] = op.direct(e.reshape(op.coeff_shape)).ravel()
        e[j] = 0.0
    return A


def test_dual_prox_optimality():
    """prox_{s F*}(v) = v - s prox_{F/s}(v/s) (Moreau); prox_{F/s}(w) = u solves sc (1 - b/(u + r))/s + u - w = 0."""
    rng = np.random.default_rng(0)
    n = 1000
    y, Ax = rng.normal(size=n), rng.random(n) * 10
    s, b = rng.random(n) + 0.1, rng.poisson(5, n).astype(float)
    w, g, bg = rng.random(n) + 0.5, rng.random(n) + 0.5, rng.random(n)
    eps = 1e-3
    yn = O.dual_prox(y, Ax, s, b, w, g, bg, eps)
    v = y + s * Ax
    u = (v - yn) / s                                  # prox_{F/s}(v/s)
    sc, r = w / g, bg + eps * g
    np.testing.assert_allclose(sc * (1 - b / (u + r)) / s + u - v / s, 0.0, atol=1e-9)
    assert np.all(u + r > 0)


def test_spdhg_reaches_the_minimiser():
    op = _op()
    A = _matrix(op)
    rng = np.random.default_rng(7)
    supp = op.support_mask()
    xt = (rng.random(op.coeff_shape) * (rng.random(op.coeff_shape) < 0.3) * supp[None])
    gain = 0.05
    b = rng.poisson(op.direct(xt).clip(0) / gain) * gain
    eps = float(np.float32(1e-3))

    def f(x):
        mu = np.maximum(A @ x, 0) + eps * gain
        bb = b.ravel()
        v = mu - bb + np.where(bb > 0, bb * np.log(np.where(bb > 0, bb, 1) / mu), 0)
        return float(v.sum() / gain)

    def fg(x):
        mu = A @ x + eps * gain
        return f(x), A.T @ ((1 - b.ravel() / mu) / gain)

    free = np.broadcast_to(supp, op.coeff_shape).ravel() & (A.sum(axis=0) > 0)
    bounds = [(0, None) if fr else (0, 0) for fr in free]
    ref = minimize(fg, np.full(A.shape[1], 0.1) * free, jac=True, method="L-BFGS-B", bounds=bounds,
                   options={"maxiter": 20000, "ftol": 1e-15, "gtol": 1e-12})
    sp = O.OracleSPDHG(op, subsets=4, seed=0)
    x = sp.run(np.zeros(op.coeff_shape), b, 600, gain=gain, diagnostics_interval=600)
    assert abs(sp.final_stats["final_f"] - f(x.ravel())) <= 1e-9 * f(x.ravel())
    assert sp.final_stats["final_f"] <= ref.fun * (1 + 2e-3)


def test_power_iteration_is_the_largest_eigenvalue():
    op = _op(K=3)
    A = _matrix(op)
    lam = np.linalg.eigvalsh(A.T @ A)[-1]
    L = O.estimate_L_power(op, niter=400, seed=1)
    assert abs(L / lam - 1) < 1e-6
