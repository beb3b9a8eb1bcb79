"""Angular quadrature of the pole density over one (omega, eta) data bin, with an a-priori error bound.

A PF entry integrates g(v(omega, eta)) = exp(-(1 - |U h . v|) / sigma^2) over the bin. Along either angle the
integrand is at worst exp(a (cos t - 1)) with a = 1/sigma^2 (a great circle through the pole), whose m-th
derivative is largest at t = 0. A composite Gauss-Legendre rule of k panels of p points on a width w then errs by
at most

    E(p, k) = c_p (w / k)^(2p) * D_2p(a),       c_p = (p!)^4 / ((2p + 1) ((2p)!)^3),

with D_m(a) = max_t |d^m/dt^m exp(a (cos t - 1))|. Writing exp(a cos t) = sum_n I_n(a) e^{i n t}, D_m(a) is the
m-th moment of the symmetric Skellam law P(n) = e^{-a} I_n(a), whose cumulants are a (even) and 0 (odd); its
moment generating function is exp(a (cosh s - 1)), so D_2p(a) = sum_j B(2p, j) a^j with B(n, j) the number of
partitions of n items into j blocks of even size. The coefficients are integers, built once exactly below.

The rule of a bin is the pair of per-direction rules (p_omega, k_omega, p_eta, k_eta) with the fewest nodes
p_omega k_omega p_eta k_eta whose summed bound leaves room for the Gaussian's cut (exp(-PF_CUT), the share of a
PF entry the cut can drop) within the tolerance; among equal node counts the smaller bound, then fewer omega
nodes. This is the reference's rule (operators/quadrature.py:127-153, 168-222): the same bound and the same choice,
so the same nodes; the search here is one vectorised pass over the table of per-direction bounds.
"""
from __future__ import annotations

from dataclasses import dataclass
from fractions import Fraction
from functools import lru_cache
from math import comb, factorial

import numpy as np

DEFAULT_TOL = 0.05      # bound on a PF entry's quadrature error relative to the peak (operator's quadrature_tol)
MAX_POINTS = 8          # Gauss points per panel considered
MAX_NODES = 64          # nodes per bin and direction considered


# ------------------------------------------------------------------------------------------------- error bound
def gauss_constant(p: int) -> float:
    """c_p of the p-point Gauss-Legendre remainder on a unit panel (1/24 for the midpoint rule)."""
    return float(Fraction(factorial(p) ** 4, (2 * p + 1) * factorial(2 * p) ** 3))


@lru_cache(maxsize=None)
def _even_block_partitions(n: int) -> tuple[int, ...]:
    """(B(n, 0), ..., B(n, n)): partitions of n labelled items into j blocks of even size. The block holding
    item 1 has an even size b (C(n - 1, b - 1) ways to fill it); the rest is a partition of n - b items."""
    if n == 0:
        return (1,)
    row = [0] * (n + 1)
    for b in range(2, n + 1, 2):
        rest = _even_block_partitions(n - b)
        ways = comb(n - 1, b - 1)
        for j, c in enumerate(rest):
            row[j + 1] += ways * c
    return tuple(row)


def peak_derivative_factor(m: int, a: float) -> float:
    """D_m(a) = max_t |d^m/dt^m exp(a (cos t - 1))| for even m: the polynomial sum_j B(m, j) a^j."""
    if m % 2:
        raise ValueError("m must be even")
    coef = _even_block_partitions(int(m))
    acc = 0.0
    for c in reversed(coef):                       # Horner in a (coefficients are exact integers)
        acc = acc * float(a) + float(c)
    return acc


def direction_bound(p: int, k: int, width: float, sigma: float) -> float:
    """E(p, k): the error bound of k panels of p Gauss points over an angular width (radians)."""
    return gauss_constant(p) * (abs(float(width)) / k) ** (2 * p) * peak_derivative_factor(2 * p, 1.0 / sigma ** 2)


def nodes_weights(p: int, k: int):
    """Nodes in (0, 1) (panel by panel) and weights (summing to 1) of k panels of p-point Gauss-Legendre."""
    g, w = np.polynomial.legendre.leggauss(p)
    x = ((np.arange(k)[:, None] + (g[None, :] + 1.0) / 2.0) / k).ravel()
    return x, np.tile(w / 2.0 / k, k)


def _direction_table(width: float, sigma: float, p_max: int, n_max: int):
    """For n = 1 .. n_max nodes along one direction: the least bound over the factorings n = p k (p <= p_max), the
    fewest points per panel on ties. Returns (bounds (n_max,), p (n_max,), k (n_max,))."""
    a = 1.0 / float(sigma) ** 2
    n = np.arange(1, n_max + 1)
    best = np.full(n_max, np.inf)
    bp = np.zeros(n_max, np.int64)
    for p in range(1, p_max + 1):
        ok = n % p == 0
        k = n[ok] // p
        e = gauss_constant(p) * (abs(float(width)) / k) ** (2 * p) * peak_derivative_factor(2 * p, a)
        better = e < best[ok]                       # strict: the smaller p stays on ties
        idx = np.flatnonzero(ok)[better]
        best[idx] = e[better]
        bp[idx] = p
    return best, bp, n // bp


@dataclass(frozen=True)
class QuadratureRule:
    """k panels of p Gauss-Legendre points along omega and along eta (p = 1: midpoint sub-bins)."""
    p_omega: int = 1
    k_omega: int = 1
    p_eta: int = 1
    k_eta: int = 1

    def __post_init__(self):
        if any(int(v) != v or v < 1 for v in (self.p_omega, self.k_omega, self.p_eta, self.k_eta)):
            raise ValueError(f"QuadratureRule needs positive integers, not {self}")

    @classmethod
    def midpoint(cls, n_omega: int = 1, n_eta: int = 1) -> "QuadratureRule":
        return cls(1, int(n_omega), 1, int(n_eta))

    @property
    def n_omega(self) -> int:
        return self.p_omega * self.k_omega

    @property
    def n_eta(self) -> int:
        return self.p_eta * self.k_eta

    @property
    def is_midpoint(self) -> bool:
        return self.p_omega == 1 and self.p_eta == 1

    def bound(self, d_omega: float, d_eta: float, sigma: float) -> float:
        """The rule's error bound relative to the peak (both directions, without the cut)."""
        return direction_bound(self.p_omega, self.k_omega, d_omega, sigma) + \
            direction_bound(self.p_eta, self.k_eta, d_eta, sigma)

    def describe(self) -> str:
        n = self.n_omega * self.n_eta
        return f"omega {self.p_omega}x{self.k_omega}, eta {self.p_eta}x{self.k_eta} ({n} node{'s' * (n > 1)} per bin)"


def choose_rule(d_omega, d_eta, sigma, tol=DEFAULT_TOL, cut_share=0.0, p_max=MAX_POINTS, n_max=MAX_NODES):
    """The fewest-node rule whose bound is within tol - cut_share (ties: smaller bound, then fewer omega nodes)."""
    budget = float(tol) - float(cut_share)
    if budget <= 0:
        raise ValueError(f"quadrature_tol = {tol:g} is not above the error of the Gaussian's cut ({cut_share:.2g} "
                         f"of the peak): raise pf_cutoff_sigma or the tolerance")
    bw, pw, kw = _direction_table(d_omega, sigma, p_max, n_max)
    be, pe, ke = _direction_table(d_eta, sigma, p_max, n_max)
    total = bw[:, None] + be[None, :]
    nodes = np.arange(1, n_max + 1)[:, None] * np.arange(1, n_max + 1)[None, :]
    feasible = np.flatnonzero((total <= budget).ravel())
    if feasible.size == 0:
        raise ValueError(f"no rule with at most {n_max} nodes per direction meets quadrature_tol = {tol:g} for bins "
                         f"of {np.degrees(abs(d_omega)):.3g} x {np.degrees(abs(d_eta)):.3g} deg and sigma = "
                         f"{np.degrees(sigma):.3g} deg: use a larger tolerance or pass a QuadratureRule")
    # lexicographic (nodes, bound, omega nodes): lexsort's last key is the primary one; flat order is omega-major
    order = np.lexsort((feasible, total.ravel()[feasible], nodes.ravel()[feasible]))
    iw, ie = np.unravel_index(feasible[order[0]], total.shape)
    return QuadratureRule(int(pw[iw]), int(kw[iw]), int(pe[ie]), int(ke[ie]))


def cut_share(pf_cutoff_sigma) -> float:
    """The PF error (relative to the peak) the cut of the Gaussian at exponent PF_CUT can cause: exp(-PF_CUT)."""
    return float(np.exp(-(6.0 if pf_cutoff_sigma is None else float(pf_cutoff_sigma) ** 2 / 2)))


def _bin_nodes(lo, hi, n_bins, p, k):
    """Node angles of every bin, bin by bin. Midpoint sub-bins use linspace over all n_bins * k sub-bins (the
    reference's expression, so the angles agree to the last bit); Gauss panels lo + (bin + x) * width."""
    width = (hi - lo) / n_bins
    if p == 1:
        return np.linspace(lo, hi, n_bins * k, endpoint=False) + width / k / 2
    x, _ = nodes_weights(p, k)
    return (lo + (np.arange(n_bins)[:, None] + x[None, :]) * width).ravel()


def setup_quadrature(op, quadrature, tol, sigma_min):
    """Resolve the operator's quadrature option ("auto", "midpoint" or a QuadratureRule) and set quadrature,
    quadrature_bound (with the cut's share), N_Omega_subdivisions, N_eta_subdivisions, angles_subdivided,
    eta_angles_subdivided and quad_weights_cpu ((S_omega S_eta,) float32 product weights; None for midpoint rules,
    whose kernels average), from op.angle_range / eta_angle_range (radians), N_Omega, N_eta, cfg and
    pf_cutoff_sigma."""
    a0, a1 = (float(v) for v in op.angle_range)
    e0, e1 = (float(v) for v in op.eta_angle_range)
    d_om, d_eta = (a1 - a0) / op.N_Omega, (e1 - e0) / op.N_eta
    share = cut_share(op.pf_cutoff_sigma)
    if isinstance(quadrature, QuadratureRule):
        rule = quadrature
    elif quadrature == "midpoint":
        rule = QuadratureRule.midpoint(op.cfg.get("N_Omega_subdivisions", 1), op.cfg.get("N_eta_subdivisions", 1))
    elif quadrature == "auto":
        rule = choose_rule(d_om, d_eta, sigma_min, tol, share)
    else:
        raise ValueError(f"quadrature must be 'auto', 'midpoint' or a QuadratureRule, not {quadrature!r}")
    op.quadrature = rule
    op.quadrature_bound = rule.bound(d_om, d_eta, sigma_min) + share
    op.N_Omega_subdivisions, op.N_eta_subdivisions = rule.n_omega, rule.n_eta
    op.angles_subdivided = _bin_nodes(a0, a1, op.N_Omega, rule.p_omega, rule.k_omega)
    op.eta_angles_subdivided = _bin_nodes(e0, e1, op.N_eta, rule.p_eta, rule.k_eta)
    if rule.is_midpoint:
        op.quad_weights_cpu = None
    else:
        ww = nodes_weights(rule.p_omega, rule.k_omega)[1]
        we = nodes_weights(rule.p_eta, rule.k_eta)[1]
        op.quad_weights_cpu = np.ascontiguousarray(np.outer(ww, we).ravel(), dtype=np.float32)
    return rule
