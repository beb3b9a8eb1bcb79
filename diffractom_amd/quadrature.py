"""Quadrature of the pole density over one (omega, eta) data bin with an error bound (the reference's
operators/quadrature.py): tensor-product Gauss-Legendre, k panels of p points per direction, the fewest nodes whose
bound c_p h^(2p) mu_2p(1/sigma^2) (summed over both directions, plus the share of the Gaussian's cut) meets the
tolerance."""
from __future__ import annotations

from dataclasses import dataclass
from math import comb, factorial

import numpy as np

P_MAX = 8
N_MAX = 64
DEFAULT_TOL = 0.05


def gauss_constant(p: int) -> float:
    """(p!)^4 / ((2p + 1) ((2p)!)^3): 1/24 for the midpoint rule."""
    return factorial(p) ** 4 / ((2 * p + 1) * factorial(2 * p) ** 3)


def peak_derivative_factor(m: int, a: float) -> float:
    """Largest |d^m/dt^m exp(a (cos t - 1))| for even m: the m-th moment of the symmetric Skellam distribution
    (cumulants a for even orders, 0 for odd), by the moment-cumulant recursion."""
    if m % 2:
        raise ValueError("m must be even")
    mom = [1.0]
    for n in range(1, m + 1):
        mom.append(sum(comb(n - 1, j - 1) * a * mom[n - j] for j in range(2, n + 1, 2)))
    return mom[m]


def direction_bound(p: int, k: int, width: float, sigma: float) -> float:
    h = abs(float(width)) / k
    return gauss_constant(p) * h ** (2 * p) * peak_derivative_factor(2 * p, 1.0 / float(sigma) ** 2)


def nodes_weights(p: int, k: int):
    """Node offsets in (0, 1) and weights (sum 1) of k panels of p-point Gauss-Legendre."""
    g, w = np.polynomial.legendre.leggauss(p)
    x = ((np.arange(k)[:, None] + (g[None, :] + 1.0) / 2.0) / k).ravel()
    return x, np.tile(w / 2.0 / k, k)


@dataclass(frozen=True)
class QuadratureRule:
    p_omega: int = 1
    k_omega: int = 1
    p_eta: int = 1
    k_eta: int = 1

    def __post_init__(self):
        for v in (self.p_omega, self.k_omega, self.p_eta, self.k_eta):
            if int(v) != v or v < 1:
                raise ValueError(f"QuadratureRule needs positive integers, not {self}")

    @classmethod
    def midpoint(cls, n_omega: int = 1, n_eta: int = 1):
        return cls(1, int(n_omega), 1, int(n_eta))

    @property
    def n_omega(self):
        return self.p_omega * self.k_omega

    @property
    def n_eta(self):
        return self.p_eta * self.k_eta

    @property
    def is_midpoint(self):
        return self.p_omega == 1 and self.p_eta == 1

    def bound(self, d_omega, d_eta, sigma):
        return direction_bound(self.p_omega, self.k_omega, d_omega, sigma) + \
            direction_bound(self.p_eta, self.k_eta, d_eta, sigma)

    def describe(self):
        return f"omega {self.p_omega}x{self.k_omega}, eta {self.p_eta}x{self.k_eta} " \
               f"({self.n_omega * self.n_eta} nodes per bin)"


def _options(width, sigma, p_max, n_max):
    out = []
    for n in range(1, n_max + 1):
        best = None
        for p in range(1, min(p_max, n) + 1):
            if n % p:
                continue
            b = direction_bound(p, n // p, width, sigma)
            if best is None or b < best[0]:
                best = (b, p, n // p)
        out.append((n, *best))
    return out


def choose_rule(d_omega, d_eta, sigma, tol=DEFAULT_TOL, cut_share=0.0, p_max=P_MAX, n_max=N_MAX):
    budget = float(tol) - float(cut_share)
    if budget <= 0:
        raise ValueError(f"quadrature_tol = {tol:g} is not above the error of the Gaussian's cut ({cut_share:.2g})")
    best = None
    for nw, bw, pw, kw in _options(d_omega, sigma, p_max, n_max):
        for ne, be, pe, ke in _options(d_eta, sigma, p_max, n_max):
            if bw + be <= budget:
                key = (nw * ne, bw + be)
                if best is None or key < best[0]:
                    best = (key, QuadratureRule(pw, kw, pe, ke))
                break
    if best is None:
        raise ValueError(f"no rule with at most {n_max} nodes per direction meets quadrature_tol = {tol:g}")
    return best[1]


def cut_share(pf_cutoff_sigma) -> float:
    return float(np.exp(-(6.0 if pf_cutoff_sigma is None else float(pf_cutoff_sigma) ** 2 / 2)))


def setup_quadrature(op, quadrature, tol, sigma_min):
    """Sets op.quadrature, quadrature_bound, N_Omega_subdivisions, N_eta_subdivisions, angles_subdivided,
    eta_angles_subdivided and quad_weights_cpu (None for the midpoint rule) from op.angle_range / eta_angle_range
    (radians), op.N_Omega, op.N_eta, op.cfg and op.pf_cutoff_sigma (quadrature.py:168-222)."""
    cfg = op.cfg
    a0, a1 = float(op.angle_range[0]), float(op.angle_range[1])
    e0, e1 = float(op.eta_angle_range[0]), float(op.eta_angle_range[1])
    delta = (a1 - a0) / op.N_Omega
    eta_delta = (e1 - e0) / op.N_eta
    share = cut_share(op.pf_cutoff_sigma)
    if isinstance(quadrature, QuadratureRule):
        rule = quadrature
    elif quadrature == "midpoint":
        rule = QuadratureRule.midpoint(cfg.get("N_Omega_subdivisions", 1), cfg.get("N_eta_subdivisions", 1))
    elif quadrature == "auto":
        rule = choose_rule(delta, eta_delta, sigma_min, tol, share)
    else:
        raise ValueError(f"quadrature must be 'auto', 'midpoint' or a QuadratureRule, not {quadrature!r}")
    op.quadrature = rule
    op.quadrature_bound = rule.bound(delta, eta_delta, sigma_min) + share
    S, Se = rule.n_omega, rule.n_eta
    op.N_Omega_subdivisions, op.N_eta_subdivisions = S, Se
    if rule.p_omega == 1:
        op.angles_subdivided = np.linspace(a0, a1, op.N_Omega * S, endpoint=False) + delta / S / 2
    else:
        x, _ = nodes_weights(rule.p_omega, rule.k_omega)
        op.angles_subdivided = (a0 + (np.arange(op.N_Omega)[:, None] + x[None, :]) * delta).ravel()
    if rule.p_eta == 1:
        op.eta_angles_subdivided = np.linspace(e0, e1, op.N_eta * Se, endpoint=False) + eta_delta / Se / 2
    else:
        x, _ = nodes_weights(rule.p_eta, rule.k_eta)
        op.eta_angles_subdivided = (e0 + (np.arange(op.N_eta)[:, None] + x[None, :]) * eta_delta).ravel()
    if rule.is_midpoint:
        op.quad_weights_cpu = None
    else:
        _, ww = nodes_weights(rule.p_omega, rule.k_omega)
        _, we = nodes_weights(rule.p_eta, rule.k_eta)
        op.quad_weights_cpu = np.ascontiguousarray((ww[:, None] * we[None, :]).ravel(), dtype=np.float32)
    return rule
