"""The FP64 differentiable model (oracle/diffmodel.py) pinned independently of its own formulas: at the initial
parameters it reproduces the oracle operator; every analytic derivative equals central finite differences of the
model (no cut: smooth), at parameters moved away from the initial ones (the tilt at 0, where its increment and its
parameter derivatives coincide), for the trapezoid and the hat footprint
(the reference's own check, tests/test_differentiable.py:38-66); the Gauss-Newton blocks equal J^T diag(curv) J
with J from finite differences."""
import numpy as np
import pytest
from scipy.spatial.transform import Rotation

from oracle.diffmodel import OracleDiffModel, perturb
from oracle.model import OracleOperator
from pairs import AL3, cfg, orientations


def _oracle(kernel="trapezoid", quadrature="auto", K=8):
    c = cfg(nx=12, ny=10, my=15, n_omega=10, angle_range=(0, 360), cor=0.7, n_eta=16)
    U = orientations(K, 5)
    return OracleOperator(c, (4.05, 4.05, 4.05, 90, 90, 90), "cubic", AL3, 0.35, U, np.deg2rad(6.0),
                          quadrature=quadrature, radon_kernel=kernel, orthonormalize=False)


def _moved_rot(prm, k, axis, h):
    out = perturb(prm, "rot", (k, axis), 0.0)
    e = np.zeros(3)
    e[axis] = h
    out["rot"][k] = (Rotation.from_rotvec(e) * Rotation.from_rotvec(prm["rot"][k])).as_rotvec()
    return out


def test_model_reproduces_the_oracle_operator():
    orc = _oracle()
    m = OracleDiffModel(orc)
    rng = np.random.default_rng(0)
    c = rng.random(orc.coeff_shape)
    d = m.forward(c, m.params0())
    assert np.abs(d - orc.direct(c)).max() < 1e-12 * np.abs(d).max()


@pytest.mark.parametrize("kernel,quadrature", [("trapezoid", "auto"), ("hat", (1, 1, 1, 1))])
def test_derivatives_equal_finite_differences(kernel, quadrature):
    orc = _oracle(kernel, quadrature)
    m = OracleDiffModel(orc, cut=None)
    rng = np.random.default_rng(1)
    c = rng.random(orc.coeff_shape) * orc.support_mask()[None]
    r = rng.normal(size=orc.data_shape)
    prm = m.params0()
    prm["rot"] = rng.normal(0, 0.01, (orc.K, 3))
    prm["sigma"] = prm["sigma"] * (1 + 0.05 * rng.random(orc.K))
    prm["two_theta"] = prm["two_theta"] + 0.002
    prm["eta_offset"] = 0.01
    # the tilt derivative is that of a tilt increment on top of the current axis (the product's convention,
    # reference update.py:30-37), which is the parameter derivative at zero tilt: the tilt stays 0 here
    prm["cor"] = prm["cor"] + 0.1
    prm["shift"] = rng.normal(0, 0.2, orc.R)
    g = m.vjp(c, prm, r)
    L = lambda p_: float(np.sum(r * m.forward(c, p_)))
    checks = [("rot", (3, 1), 1e-6), ("rot", (0, 2), 1e-6), ("sigma", (5,), 1e-7), ("two_theta", (1,), 1e-6),
              ("eta_offset", None, 1e-6), ("axis_tilt", (0,), 1e-6), ("axis_tilt", (1,), 1e-6),
              ("cor", None, 1e-5), ("shift", (7,), 1e-5)]
    for grp, ix, h in checks:
        if grp == "rot":
            moved = lambda hh, ix=ix: _moved_rot(prm, ix[0], ix[1], hh)
        else:
            moved = lambda hh, grp=grp, ix=ix: perturb(prm, grp, ix, hh)
        fd = (L(moved(h)) - L(moved(-h))) / (2 * h)
        an = float(g[grp] if ix is None else g[grp][ix])
        assert abs(an - fd) <= 1e-6 * max(abs(fd), 1e-3 * np.abs(r).sum()), (grp, ix, an, fd)
    # the coefficient gradient is the adjoint: <g_c, x> = <r, A x>
    x = rng.random(orc.coeff_shape)
    assert abs(np.sum(g["c"] * x) / np.sum(r * m.forward(x, prm)) - 1) < 1e-10


def test_gauss_newton_blocks():
    orc = _oracle(K=4)
    m = OracleDiffModel(orc, cut=None)
    rng = np.random.default_rng(2)
    prm = m.params0()
    prm["rot"] = rng.normal(0, 0.01, (orc.K, 3))
    c = rng.random(orc.coeff_shape) * orc.support_mask()[None]
    curv = rng.random(orc.data_shape)
    Pm, _ = m.projection(prm)
    S = np.einsum("asp,kp->ask", Pm, c.reshape(orc.K, -1))
    H = m.pf_jacobian_blocks(prm, curv, S)
    for k in (1, 3):
        ck = np.zeros_like(c)
        ck[k] = c[k]
        cols = []
        for i in range(4):
            h = 1e-6
            if i < 3:
                fp, fm = m.forward(ck, _moved_rot(prm, k, i, h)), m.forward(ck, _moved_rot(prm, k, i, -h))
            else:
                fp, fm = m.forward(ck, perturb(prm, "sigma", (k,), h)), m.forward(ck, perturb(prm, "sigma", (k,), -h))
            cols.append(((fp - fm) / (2 * h)).ravel())
        J = np.stack(cols, 1)
        Hfd = J.T @ (curv.ravel()[:, None] * J)
        assert np.abs(H[k] - Hfd).max() <= 1e-6 * np.abs(Hfd).max(), k
