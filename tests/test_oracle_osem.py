"""The oracle's OSEM / MLEM against mathematics: the EM ascent property (MLEM never increases the deviance), the EM
count identity (without background an MLEM step makes the fitted total prediction equal the data total, up to the
eps_counts floor), the textbook MLEM / OSEM iterates on the explicitly probed matrix (Shepp and Vardi 1982; Hudson and
Larkin 1994), the reference's own bar (tests/test_rock_pipeline.py:55-83: the deviance non-increasing between
checkpoints, x >= 0, 0 off the support), and held-out angles ignored (tests/test_rock_pipeline.py:86-103)."""
import numpy as np

from oracle import model as O
from pairs import AL3, cfg, orientations


def _op(K=4):
    c = cfg(nx=8, ny=8, my=9, n_omega=8, angle_range=(0, 180), cor=0.0, n_eta=12)
    return O.OracleOperator(c, (4.05, 4.05, 4.05, 90, 90, 90), "cubic", AL3, 0.35, orientations(K, 5),
                            np.deg2rad(6.0), orthonormalize=False)


def _data(op, seed=7, gain=1.0, scale=50.0):
    rng = np.random.default_rng(seed)
    supp = op.support_mask()
    xt = rng.random(op.coeff_shape) * (rng.random(op.coeff_shape) < 0.3) * supp[None]
    ax = op.direct(xt).clip(0)
    return rng.poisson(ax / ax.mean() * scale / gain) * gain


def test_mlem_never_increases_the_deviance():
    op = _op()
    b = _data(op)
    em = O.OracleOSEM(op, subsets=None)
    em.run(np.zeros(op.coeff_shape), b, 30, diagnostics_interval=1)
    f = np.array([s["f"] for s in em.iter_stats])
    assert np.all(np.diff(f) <= 1e-12 * f[0])
    assert f[-1] < 0.5 * f[0]


def test_mlem_step_matches_the_fitted_total():
    """sum_i (A x')_i = sum_j (A^T 1)_j x'_j = sum_i b_i (A x)_i / mu_i: the data total up to eps / mu."""
    op = _op()
    b = _data(op)
    em = O.OracleOSEM(op, subsets=None)
    x0 = np.random.default_rng(3).random(op.coeff_shape) + 0.1
    for n in (1, 2, 5):
        x = em.run(x0, b, n, start_floor=0, diagnostics_interval=0)
        assert abs(op.direct(x).sum() / b.sum() - 1) < 1e-4


def _matrix(op):
    n = int(np.prod(op.coeff_shape))
    A = np.zeros((int(np.prod(op.data_shape)), n))
    e = np.zeros(n)
    for j in range(n):
        e[j] = 1.0
        A[:, j] = op.direct(e.reshape(op.coeff_shape)).ravel()
        e[j] = 0.0
    return A


def test_em_iterates_are_the_textbook_ones_on_the_explicit_matrix():
    """Shepp-Vardi MLEM and Hudson-Larkin OSEM written on the probed matrix A (rows of subset m: its angles'
    points): x <- x A_m^T(b / mu) / A_m^T 1 with mu = A x + eps g, the gain cancelling (w = 1), x = 0 off the
    support; the flat start c = sum(b) / sum_allowed(A^T 1); subset order default_rng(seed).permutation(M)."""
    op = _op(K=3)
    A = _matrix(op)
    supp = np.broadcast_to(op.support_mask(), op.coeff_shape).ravel()
    gain = 0.05
    b = _data(op, seed=5, gain=gain).ravel()
    eps = float(np.float32(1e-3))
    row = int(np.prod(op.data_shape[1:]))
    for M, seed, lam in ((None, 0, 0.0), (3, 2, 0.0), (4, 1, 30.0)):
        em = O.OracleOSEM(op, subsets=M, seed=seed, lam=lam)
        got = em.run(np.zeros(op.coeff_shape), b.reshape(op.data_shape), 6, gain=gain, diagnostics_interval=0)
        subs = [np.arange(op.data_shape[0])] if M is None else [np.arange(m, op.data_shape[0], M) for m in range(M)]
        rows = [np.concatenate([np.arange(a * row, (a + 1) * row) for a in sub]) for sub in subs]
        sens_all = A.T @ np.full(A.shape[0], 1.0 / gain)
        x = np.where(supp, bsum := float(b.sum()) / gain / float(sens_all[supp].sum()), 0.0)
        assert abs(em.start_floor_used - bsum) <= 1e-12 * bsum
        rng = np.random.default_rng(seed)
        for _ in range(6):
            order = rng.permutation(len(subs)) if len(subs) > 1 else [0]
            for m in order:
                Am = A[rows[m]]
                mu = Am @ x + eps * gain
                back = Am.T @ (b[rows[m]] / gain / mu)
                sens = Am.T @ np.full(len(rows[m]), 1.0 / gain)
                x = np.where(supp, np.where(sens > 0, x * back / np.where(sens > 0, sens + lam, 1.0), x), 0.0)
        np.testing.assert_allclose(got.ravel(), x, rtol=1e-9, atol=1e-12 * np.abs(x).max())


def test_osem_deviance_does_not_increase_between_checkpoints():
    """The reference's own bar (tests/test_rock_pipeline.py:69-75): OSEM with 4 subsets, the deviance every 50
    epochs non-increasing; x >= 0 and 0 off the support; final_f the deviance of the returned x."""
    op = _op()
    supp = op.support_mask()
    gain = 0.05
    b = _data(op, seed=11, gain=gain)
    em = O.OracleOSEM(op, subsets=4, seed=0)
    x = em.run(np.zeros(op.coeff_shape), b, 200, gain=gain, diagnostics_interval=50)
    f = [s["f"] for s in em.iter_stats if "f" in s]
    assert len(f) == 4 and all(np.diff(f) <= 1e-6 * f[0])
    assert x.min() >= 0 and (x * ~supp[None]).max() == 0
    assert em.final_f == O.deviance(op.direct(x), b, None, np.full(1, gain), None, 1e-3)


def test_held_out_angles_do_not_change_the_fit():
    op = _op()
    b = _data(op, seed=2)
    fit = np.ones(op.data_shape[0], bool)
    fit[[1, 5]] = False
    subs = [np.flatnonzero(fit)[i::3] for i in range(3)]
    b2 = b.copy()
    b2[~fit] = np.random.default_rng(0).uniform(0, 10, b2[~fit].shape)
    xs = [O.OracleOSEM(op, subsets=subs, seed=4).run(np.zeros(op.coeff_shape), bb, 10, diagnostics_interval=0)
          for bb in (b, b2)]
    np.testing.assert_array_equal(xs[0], xs[1])
