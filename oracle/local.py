"""FP64 oracle of the local-basis operator (operators/local_forward_operator.py:1-20, 394-425).

TEST INFRASTRUCTURE: imported only by tests/ as the checker; the product never imports it.

The reference defines the local operator as SinglePhaseForwardOperator applied to the coefficients scattered into
(K, Ny, Nx) (zero outside the support), and its adjoint as the global adjoint gathered at the allowed pairs (zero at
the other coefficients of the patches). This oracle restates exactly that on top of OracleOperator (pinned by the
brute-force, closed-form and notebook checks of tests/test_oracle_pf_radon.py and tests/test_oracle_golden.py): it
takes the allowed pairs (k, iy, ix) and the flat index of each pair's coefficient, the layout under test. That
layout itself is pinned on the CPU by tests/test_local_support.py (every pair in one box of its orientation at its
row-major index, boxes of an orientation disjoint, the reference's split and dilation examples), so the scatter and
gather below apply a checked map.

The SPDHG run on a local basis (spdhg_poisson.py:754-943) is, without total variation, SPDHG on (K, Ny, Nx) with the
coefficients outside the support held at zero: OracleSPDHG with a (K, Ny, Nx) support mask (with total variation
also when every orientation is one patch: its neighbours are then the image's neighbours within the support).
"""
from __future__ import annotations

import numpy as np


class OracleLocalOperator:
    """A (n_coef,) flat-coefficient operator over an OracleOperator `base` whose orientations are the basis's:
    pairs (k, iy, ix) of the support and pair_index (the flat index of every pair's coefficient)."""

    def __init__(self, base, k, iy, ix, pair_index, n_coef):
        self.base = base
        self.k, self.iy, self.ix = (np.asarray(a, np.int64) for a in (k, iy, ix))
        self.pi = np.asarray(pair_index, np.int64)
        self.n_coef = int(n_coef)

    @property
    def coeff_shape(self):
        return (self.n_coef,)

    @property
    def data_shape(self):
        return self.base.data_shape

    def scatter(self, x):
        """(K, Ny, Nx) float64: the allowed pairs' coefficients, zero elsewhere."""
        out = np.zeros(self.base.coeff_shape)
        out[self.k, self.iy, self.ix] = np.asarray(x, np.float64).reshape(-1)[self.pi]
        return out

    def gather(self, c):
        """(n_coef,) float64: c at the allowed pairs, zero at the other coefficients."""
        out = np.zeros(self.n_coef)
        out[self.pi] = np.asarray(c, np.float64)[self.k, self.iy, self.ix]
        return out

    def support_mask3(self):
        """(K, Ny, Nx) bool: the allowed pairs."""
        m = np.zeros(self.base.coeff_shape, bool)
        m[self.k, self.iy, self.ix] = True
        return m

    def set_omega_subsets(self, subsets):
        return self.base.set_omega_subsets(subsets)

    def direct(self, x, subset=None):
        return self.base.direct(self.scatter(x), subset=subset)

    def direct_rows(self, x, rows):
        return self.base.direct_rows(self.scatter(x), rows)

    def adjoint(self, y, subset=None):
        return self.gather(self.base.adjoint(y, subset=subset))
