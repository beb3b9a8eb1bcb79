"""The oracle's quadrature and the product's against the reference's own known answers
(tests/test_quadrature.py:14-92 of the reference) and the mathematics (exactness, the error bound)."""
import numpy as np
import pytest
from scipy.special import ive

from oracle import model as O
from diffractom_amd import quadrature as Q

D = np.radians
IMPLS = [("oracle", O), ("product", Q)]


@pytest.mark.parametrize("name,m", IMPLS)
@pytest.mark.parametrize("p,k", [(1, 1), (1, 4), (2, 1), (3, 2), (5, 1), (8, 3)])
def test_nodes_weights_exact_for_polynomials(name, m, p, k):
    x, w = m.nodes_weights(p, k)
    assert np.all((x > 0) & (x < 1)) and np.isclose(w.sum(), 1.0)
    for deg in range(2 * p):
        assert np.isclose((w * x ** deg).sum(), 1.0 / (deg + 1), rtol=1e-12, atol=1e-14)


@pytest.mark.parametrize("name,m", IMPLS)
def test_gauss_constant_and_moments(name, m):
    assert m.gauss_constant(1) == pytest.approx(1 / 24)
    assert m.gauss_constant(2) == pytest.approx(1 / 4320)
    for a in (0.3, 5.0, 300.0):
        assert m.peak_derivative_factor(2, a) == pytest.approx(a)
        assert m.peak_derivative_factor(4, a) == pytest.approx(3 * a ** 2 + a)
        assert m.peak_derivative_factor(6, a) == pytest.approx(15 * a ** 3 + 15 * a ** 2 + a)
        assert m.peak_derivative_factor(8, a) == pytest.approx(105 * a ** 4 + 210 * a ** 3 + 63 * a ** 2 + a)
        kk = np.arange(-600, 601)
        for mm in (2, 4, 8, 12):
            assert m.peak_derivative_factor(mm, a) == pytest.approx((kk.astype(float) ** mm * ive(kk, a)).sum(),
                                                                    rel=1e-9)


def test_peak_derivative_factor_is_the_maximum_derivative():
    t = np.linspace(-np.pi, np.pi, 200001)
    for a in (0.5, 7.0, 2000.0):
        f = np.exp(a * (np.cos(t) - 1))
        d2 = f * (a ** 2 * np.sin(t) ** 2 - a * np.cos(t))
        assert np.abs(d2).max() == pytest.approx(O.peak_derivative_factor(2, a), rel=1e-6)


def test_bound_holds_along_a_great_circle():
    rng = np.random.default_rng(0)
    s = D(0.4)
    gx, gw = np.polynomial.legendre.leggauss(64)
    for width, p, k in ((D(0.5), 1, 1), (D(0.5), 2, 1), (D(2.0), 4, 1), (D(2.0), 3, 2)):
        x, w = O.nodes_weights(p, k)
        b = O.direction_bound(p, k, width, s)
        worst = 0.0
        for u in rng.uniform(-3 * s, 3 * s, 100):
            f = lambda t: np.exp(-(1 - np.cos(u + t * width)) / s ** 2)
            ref = sum((f((j + (gx + 1) / 2) / 8) * gw / 2 / 8).sum() for j in range(8))
            worst = max(worst, abs((f(x) * w).sum() - ref))
        assert 0.02 * b <= worst <= b


def test_choose_rule_known_answers_and_minimality():
    share = float(np.exp(-8.0))
    assert O.choose_rule(D(0.5), D(2.0), D(4.0), 0.05, share) == (1, 1, 1, 1)
    assert O.choose_rule(D(0.5), D(2.0), D(0.4), 0.05, share) == (2, 1, 4, 1)      # the reference's simulations
    assert Q.choose_rule(D(0.5), D(2.0), D(0.4), 0.05, share) == Q.QuadratureRule(2, 1, 4, 1)
    for s in (0.2, 0.4, 1.0, 2.0):
        r = Q.choose_rule(D(0.5), D(1.0), D(s), 0.05, share)
        assert (r.p_omega, r.k_omega, r.p_eta, r.k_eta) == O.choose_rule(D(0.5), D(1.0), D(s), 0.05, share)
        n = r.n_omega * r.n_eta
        for p1 in range(1, 4):
            for k1 in range(1, 7):
                for p2 in range(1, 4):
                    for k2 in range(1, 7):
                        if p1 * k1 * p2 * k2 < n:
                            assert Q.QuadratureRule(p1, k1, p2, k2).bound(D(0.5), D(1.0), D(s)) + share > 0.05
    with pytest.raises(ValueError):
        O.choose_rule(D(0.5), D(1.0), D(0.4), 0.01, float(np.exp(-4.5)))


def test_nodes_per_bin_of_the_reference_benchmark():
    """benchmarks/RESULTS.md:17: 4 nodes per bin at sigma 0.4 deg, 2 at 1 deg, 1 from 1.3 deg (360 omega over
    180 deg, 360 eta)."""
    share = float(np.exp(-8.0))
    n = lambda s: np.prod([v for v in O.choose_rule(D(0.5), D(1.0), D(s), 0.05, share)][::2]) * \
        np.prod([v for v in O.choose_rule(D(0.5), D(1.0), D(s), 0.05, share)][1::2])
    assert n(0.4) == 4 and n(1.0) == 2 and n(1.3) == 1


def test_product_rule_choice_equals_oracle_over_a_sweep():
    """The product's vectorised search picks the oracle's rule (the reference's search) for every bin size,
    kernel width and cut of a sweep, including the cases where no rule fits."""
    share_of = lambda cut: float(np.exp(-cut))
    for dw in (0.25, 0.5, 1.0, 2.0, 3.0):
        for de in (0.5, 1.0, 2.0, 4.0):
            for s in (0.15, 0.2, 0.3, 0.4, 0.6, 1.0, 1.3, 2.0, 3.0, 5.0):
                for cut in (8.0, 6.0, 4.5):
                    try:
                        r = O.choose_rule(D(dw), D(de), D(s), 0.05, share_of(cut))
                    except ValueError:
                        r = None
                    try:
                        q = Q.choose_rule(D(dw), D(de), D(s), 0.05, share_of(cut))
                        q = (q.p_omega, q.k_omega, q.p_eta, q.k_eta)
                    except ValueError:
                        q = None
                    assert q == r, (dw, de, s, cut, q, r)


@pytest.mark.parametrize("p,k", [(1, 1), (1, 3), (2, 1), (3, 2), (4, 1)])
def test_product_node_angles_equal_oracle(p, k):
    """The node angles of every bin are the oracle's to the last bit (midpoint sub-bins and Gauss panels)."""
    for lo, hi, n in ((0.0, np.pi, 180), (0.0, 2 * np.pi, 360), (np.pi, 0.0, 90)):
        nodes, w = O.rule_nodes(lo, hi, n, p, k)
        assert np.array_equal(Q._bin_nodes(lo, hi, n, p, k), nodes)
        assert np.isclose(w.sum(), 1.0)
