"""The oracle's SPDHG with the total-variation block (spdhg_poisson.py:108-193, 430-502, 664-697) against an
independent float64 solver of the same problem: diagonally preconditioned Chambolle-Pock on the explicit matrix
[A; D] (D the forward differences between free pixels), as the reference's tests/test_spdhg_tv.py:54-141 does.
Bars (the reference's): objective within 1e-3 of the independent solver's, the reported objective within 1e-4 of the
objective evaluated at the returned x, x within 5 % of the independent solution's largest entry."""
import numpy as np
import pytest
import scipy.sparse as sp

from oracle.model import OracleOperator, OracleSPDHG, tv_adjoint, tv_dual, tv_value

AL3 = [[1, 1, 1], [2, 0, 0], [2, 2, 0]]


def _problem(K=4, n=8):
    from diffractom_amd.workloads import random_fz
    cfg = {"Nx": n, "Ny": n, "My": n + 1, "N_Omega": 8, "angle_range": [0, 180], "cor_offset": 0.0, "N_eta": 12,
           "eta_angle_range": [0, 360], "energy": 12.398 / 0.35, "k_direction_0": [0, 0, 1],
           "p_direction_0": [1, 0, 0], "detector_direction_origin": [0, -1, 0],
           "detector_direction_positive_90": [0, 0, -1]}
    U = random_fz(K, "cubic", 5)
    op = OracleOperator(cfg, (4.05, 4.05, 4.05, 90, 90, 90), "cubic", AL3, 0.35, U, np.deg2rad(6.0),
                        orthonormalize=False)
    rng = np.random.default_rng(3)
    yy, xx = np.mgrid[:n, :n]
    x_true = np.zeros(op.coeff_shape)
    for k in range(K):                       # two half planes per orientation (TV-friendly)
        a = rng.uniform(0, np.pi)
        x_true[k] = rng.uniform(0.5, 1.5) * ((np.cos(a) * (xx - n / 2) + np.sin(a) * (yy - n / 2)) > 0)
    x_true *= op.support_mask()[None]
    gain = 0.02
    b = rng.poisson(np.maximum(op.direct(x_true), 0) / gain) * gain
    return op, b, gain


def _matrix(op):
    n = int(np.prod(op.coeff_shape))
    A = np.zeros((int(np.prod(op.data_shape)), n))
    e = np.zeros(op.coeff_shape)
    for j in range(n):
        e.flat[j] = 1.0
        A[:, j] = op.direct(e).ravel()
        e.flat[j] = 0.0
    return A


def _diff_matrix(K, ny, nx, free):
    """Rows (2 per pixel: right, down) of the forward differences between free pixels."""
    rows, cols, vals = [], [], []
    for k in range(K):
        for iy in range(ny):
            for ix in range(nx):
                j = (k * ny + iy) * nx + ix
                for c, (dy, dx) in enumerate(((0, 1), (1, 0))):
                    jy, jx = iy + dy, ix + dx
                    if jy < ny and jx < nx:
                        j2 = (k * ny + jy) * nx + jx
                        if free[j] and free[j2]:
                            rows += [2 * j + c, 2 * j + c]
                            cols += [j, j2]
                            vals += [-1.0, 1.0]
    return sp.coo_matrix((vals, (rows, cols)), shape=(2 * K * ny * nx, K * ny * nx)).tocsr()


def _objective(A, D, x, b, gain, tv, eps=1e-3):
    mu = np.maximum(A @ x, 0.0) + eps * gain
    f = float(np.sum((mu - b + np.where(b > 0, b * np.log(np.where(b > 0, b, 1.0) / mu), 0.0)) / gain))
    g = (D @ x).reshape(-1, 2)
    return f + tv * float(np.sqrt((g ** 2).sum(axis=1)).sum())


def _chambolle_pock(A, D, b, gain, tv, free, eps=1e-3, n_iter=20000):
    """Diagonally preconditioned primal-dual (Pock & Chambolle 2011, alpha = 1) on K = [A; D], float64."""
    Km = sp.vstack([sp.csr_matrix(A), D]).tocsr()
    absK = abs(Km)
    sig = 1.0 / np.maximum(np.asarray(absK.sum(axis=1)).ravel(), 1e-30)
    tau = 1.0 / np.maximum(np.asarray(absK.sum(axis=0)).ravel(), 1e-30)
    tau[~free] = 0.0
    nA = A.shape[0]
    x = np.zeros(A.shape[1])
    xb = x.copy()
    y = np.zeros(Km.shape[0])
    r, sc = eps * gain, 1.0 / gain
    for _ in range(n_iter):
        v = y + sig * (Km @ xb)
        va = v[:nA] + sig[:nA] * r       # prox of sigma F*, F(u) = sc (u + r - b log(u + r))
        y[:nA] = 0.5 * (va + sc - np.sqrt((va - sc) ** 2 + 4 * sig[:nA] * sc * b))
        q = v[nA:].reshape(-1, 2)
        y[nA:] = (q / np.maximum(1.0, np.sqrt((q ** 2).sum(axis=1)) / tv)[:, None]).ravel()
        xn = np.maximum(0.0, x - tau * (Km.T @ y))
        xb = 2 * xn - x
        x = xn
    return x


def test_tv_pieces_are_the_difference_operator():
    """<D x, q> = <x, D^T q> for the dual step's differences; the dual step projects onto |q| <= lam; the TV value
    is the l2 norm of the free differences."""
    rng = np.random.default_rng(0)
    K, ny, nx = 3, 5, 6
    T = (rng.random((K, ny, nx)) > 0.2).astype(float)
    x = rng.standard_normal((K, ny, nx))
    qx, qy = rng.standard_normal((2, K, ny, nx))
    free = T.ravel() > 0
    D = _diff_matrix(K, ny, nx, free)
    g = (D @ x.ravel()).reshape(K, ny, nx, 2)
    # a huge lam and step ps: q' = q + s D x exactly on the free differences
    lam, ps = 1e30, 4.0
    nqx, nqy = tv_dual(x, T, np.zeros_like(x), np.zeros_like(x), lam, ps)
    Tr = np.zeros_like(T)
    Td = np.zeros_like(T)
    Tr[:, :, :-1] = T[:, :, 1:]
    Td[:, :-1, :] = T[:, 1:, :]
    vr = np.zeros(T.shape, bool)
    vd = np.zeros(T.shape, bool)
    vr[:, :, :-1] = (T[:, :, :-1] > 0) & (T[:, :, 1:] > 0)
    vd[:, :-1, :] = (T[:, :-1, :] > 0) & (T[:, 1:, :] > 0)
    m = np.maximum(np.where(vr, T + Tr, 0), np.where(vd, T + Td, 0))
    s = np.where(m > 0, ps / (4 * np.where(m > 0, m, 1)), 0)
    np.testing.assert_allclose(nqx, s * g[..., 0], atol=1e-12)
    np.testing.assert_allclose(nqy, s * g[..., 1], atol=1e-12)
    # adjoint
    q = np.stack([qx * vr, qy * vd], -1)
    lhs = float((D @ x.ravel()) @ q.ravel())
    rhs = float(x.ravel() @ tv_adjoint(qx * vr, qy * vd).ravel())
    assert abs(lhs - rhs) < 1e-10 * max(1, abs(lhs))
    # projection and value
    px, py = tv_dual(x, T, qx, qy, 0.3, ps)
    assert np.sqrt(px ** 2 + py ** 2).max() <= 0.3 + 1e-12
    assert abs(tv_value(x, T) - np.sqrt((g ** 2).sum(-1)).sum()) < 1e-10


@pytest.mark.parametrize("tv", [0.0, 2.0])
def test_spdhg_tv_reaches_independent_solver(tv):
    op, b, gain = _problem()
    A = _matrix(op)
    K, ny, nx = op.coeff_shape
    free = (np.broadcast_to(op.support_mask()[None], op.coeff_shape).ravel() > 0) & (A.sum(axis=0) > 0)
    D = _diff_matrix(K, ny, nx, free)
    bb = b.ravel().astype(np.float64)
    x_ref = _chambolle_pock(A, D, bb, gain, max(tv, 1e-30), free)
    f_ref = _objective(A, D, x_ref, bb, gain, tv)
    sol = OracleSPDHG(op, subsets=4, tv=tv, seed=1)
    x = sol.run(np.zeros(op.coeff_shape), b, 1000, gain=gain, diagnostics_interval=1000)
    f = _objective(A, D, x.ravel(), bb, gain, tv)
    assert f <= f_ref * (1 + 1e-3), (tv, f, f_ref)
    assert abs(sol.final_stats["final_f"] - f) <= 1e-4 * f
    if tv > 0:
        assert np.abs(x.ravel() - x_ref).max() <= 0.05 * np.abs(x_ref).max()
