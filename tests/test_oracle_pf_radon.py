"""The oracle's pole figure and projector against brute force and invariants (CPU)."""
import numpy as np
import pytest

from oracle import model as O
from pairs import AL3, cfg, orientations


def _op(**kw):
    c = cfg(**kw)
    return O.OracleOperator(c, (4.05, 4.05, 4.05, 90, 90, 90), "cubic", AL3, 0.35, orientations(6, 3), np.deg2rad(4.0),
                            orthonormalize=False)


def test_pf_equals_the_sum_over_all_symmetry_images():
    """create_pfo_matrix.py:100-106: the merged pole axes weighted by their image counts give the sum over all |G|
    images S_g h (the PF summand is even in the sign of the axis). Brute force: every image, every node."""
    op = _op(n_omega=6, n_eta=12)
    G = O.cubic_rotations()
    hn = op.refl.h_vecs / np.linalg.norm(op.refl.h_vecs, axis=1, keepdims=True)
    qw = op.qw
    for r in (0, 3):
        pf = op.pf_angle(r)
        bf = np.zeros_like(pf)
        for k in range(op.K):
            s = 1.0 / op.sigma[k] ** 2
            norm = 1.0 / (8 * np.pi * op.sigma[k] ** 2)
            for p, ring in enumerate(op.rings):
                n = (op.U[k] @ (G @ hn[ring[0]]).T).T                       # (|G|, 3) all images
                for c in range(op.C):
                    tot = 0.0
                    for si in range(op.S):
                        for se in range(op.Se):
                            d = np.abs(n @ op.v[r, si, c, se, p])
                            x1 = (1 - d) * s
                            tot += qw[si, se] * np.where(x1 < op.cut, np.exp(-x1), 0.0).sum()
                    bf[k, c, p] = norm * tot
        # the merged axes are rounded to 6 decimals and renormalised (single_phase_forward_operator.py:791-795)
        assert np.abs(pf - bf).max() <= 2e-4 * np.abs(bf).max()
        assert np.count_nonzero(pf) > 0


def test_pf_cut_and_peak():
    """Entries vanish beyond the cut angle sqrt(2 cut) sigma of every pole; a pole exactly on a single probed
    direction gives count * (1 + exp(-2/sigma^2) if kept) * norm."""
    op = _op(n_omega=6, n_eta=12)
    pf = np.stack([op.pf_angle(r) for r in range(op.R)])
    assert (pf >= 0).all()
    # the angle of every probed direction to its nearest pole, per entry: nonzero only if below the cut
    for r in (1, 4):
        for p in range(op.P):
            a0, a1 = op.start[p], op.start[p + 1]
            v = op.v[r, :, :, :, p].reshape(-1, 3)
            d = np.abs(v @ op.poles[:, a0:a1].reshape(-1, 3).T).reshape(op.S, op.C, op.Se, op.K, a1 - a0)
            near = d.max(axis=(-1,)).max(axis=0).max(axis=1)            # (C, K)
            far = (1 - near) / op.sigma[None, :] ** 2 >= op.cut
            assert np.all(pf[r][:, :, p].T[far] == 0.0)


def _dense_radon(op, r):
    """differentiable/reference.py:225-238: every bin for every pixel."""
    yy, xx = np.meshgrid(np.arange(op.Ny), np.arange(op.Nx), indexing="ij")
    px, py = xx.ravel().astype(float), yy.ravel().astype(float)
    A = max(abs(op.X[r]), abs(op.Y[r]))
    t = op.X[r] * px + op.Y[r] * py + op.T0[r]
    M = np.zeros((op.My, op.Nx * op.Ny))
    for s in range(op.My):
        M[s] = op.scale * op._kernel(t - s, A)
    return M


@pytest.mark.parametrize("kernel", ["trapezoid", "hat"])
def test_radon_matrix_equals_brute_force(kernel):
    op = _op(nx=14, ny=12, my=17, n_omega=16, angle_range=(0, 360), cor=0.7)
    op.kernel = kernel
    op._P = {}
    for r in range(op.R):
        np.testing.assert_allclose(op.radon_matrix(r).toarray(), _dense_radon(op, r), rtol=0, atol=1e-15)


def test_trapezoid_row_sums_are_constant():
    """radon_kernels.cl:11-16: with the trapezoid every bin gets the same total weight from a full row of pixels at
    every angle (the hat does not); checked on bins the image covers fully."""
    op = _op(nx=64, ny=64, my=40, n_omega=13, angle_range=(0, 180), cor=0.0)
    yy, xx = np.meshgrid(np.arange(op.Ny), np.arange(op.Nx), indexing="ij")
    spreads = {}
    for kernel in ("trapezoid", "hat"):
        op.kernel = kernel
        op._P = {}
        worst = 0.0
        for r in range(op.R):
            A = max(abs(op.X[r]), abs(op.Y[r]))
            rows_y = abs(op.X[r]) >= abs(op.Y[r])
            M = op.radon_matrix(r).toarray()
            line = (yy == op.Ny // 2) if rows_y else (xx == op.Nx // 2)   # one full row along the dense axis
            sums = M[:, line.ravel()].sum(axis=1) / op.scale
            inner = sums[10:30]
            worst = max(worst, inner.max() - inner.min())
        spreads[kernel] = worst
    assert spreads["trapezoid"] < 1e-12
    assert spreads["hat"] > 1e-3


def test_hat_equals_trapezoid_at_0_and_90_degrees():
    op = _op(nx=10, ny=10, my=11, n_omega=2, angle_range=(-45, 135), cor=0.0)     # angles 0 and 90 degrees
    mats = {}
    for kernel in ("trapezoid", "hat"):
        op.kernel = kernel
        op._P = {}
        mats[kernel] = [op.radon_matrix(r).toarray() for r in range(2)]
    for a, b in zip(mats["trapezoid"], mats["hat"]):
        np.testing.assert_allclose(a, b, atol=1e-12)


def test_operator_adjoint_is_transpose_and_linear():
    op = _op()
    op.set_omega_subsets(3)
    rng = np.random.default_rng(4)
    x, y = rng.random(op.coeff_shape), rng.random(op.data_shape)
    assert abs(np.vdot(op.direct(x), y) / np.vdot(x, op.adjoint(y)) - 1) < 1e-12
    x2 = rng.random(op.coeff_shape)
    np.testing.assert_allclose(op.direct(2 * x + x2), 2 * op.direct(x) + op.direct(x2), rtol=1e-12, atol=1e-10)
    tot = sum(op.adjoint(y, subset=m) for m in range(3))
    np.testing.assert_allclose(tot, op.adjoint(y), rtol=1e-12, atol=1e-10)
