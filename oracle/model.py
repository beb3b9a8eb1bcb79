"""FP64 NumPy oracle of diffractom's SinglePhaseForwardOperator (native projector) and SPDHGPoisson.

TEST INFRASTRUCTURE: imported only by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg, as the
checker. The product (diffractom_amd) never imports it.

Every function restates the reference's algorithm in float64 and cites the reference file:line it follows
(paths relative to the diffractom source tree src/diffractom/). What pins each function is listed in its docstring
(the test that checks it); a function with no such check would say "parity unpinned".

Model (differentiable/reference.py:1-26):
    d[a, s, j] = sum_k PF[a, k, j] S_k[a, s],   S_k[a, s] = scale sum_pix K(t_a(pix) - s) c_k(pix),   j = c P + p
    PF[a, k, (c, p)] = I_p N(sigma_k) sum_nodes w_n sum_{axes h of ring p} m_h g(U_k h . v_n; sigma_k)
    g(d; sigma) = exp(-(1 - |d|)/sigma^2) [+ exp(-(1 + |d|)/sigma^2)], each term dropped where its exponent >= cut
"""
from __future__ import annotations

import itertools
import os
from dataclasses import dataclass
from math import comb, factorial

import numpy as np
import scipy.sparse as sp
from scipy.spatial.transform import Rotation

KEV_ANGSTROM = 12.398      # operators/single_phase_forward_operator.py:554 (the operator's energy <-> wavelength)


# ============================================================================ crystallography
def lattice_matrix(a, b, c, alpha, beta, gamma):
    """Direct lattice matrix, columns = lattice vectors (crystallography/material.py:72-105): a along x for a cell
    with all angles 90 deg, else the Busing-Levy frame with a* along x and c along z.
    Pinned by tests/test_oracle_crystal.py (notebook d-spacings and two-theta; hexagonal d-spacing closed form)."""
    if alpha == 90.0 and beta == 90.0 and gamma == 90.0:
        return np.diag([float(a), float(b), float(c)])
    ca, cb, cg = (np.cos(np.deg2rad(t)) for t in (alpha, beta, gamma))
    G = np.array([[a * a, a * b * cg, a * c * cb], [a * b * cg, b * b, b * c * ca], [a * c * cb, b * c * ca, c * c]])
    Gi = np.linalg.inv(G)
    ast, bst, cst = np.sqrt(np.diag(Gi))
    cos_bs = Gi[0, 2] / (ast * cst)
    cos_gs = Gi[0, 1] / (ast * bst)
    B = np.array([[ast, bst * cos_gs, cst * cos_bs],
                  [0.0, bst * np.sqrt(1 - cos_gs ** 2), -cst * np.sqrt(1 - cos_bs ** 2) * ca],
                  [0.0, 0.0, 1.0 / c]])
    return np.linalg.inv(B).T


def reciprocal(A):
    """B = 2 pi (A^-1)^T (crystallography/lattice.py:17-19)."""
    return 2.0 * np.pi * np.linalg.inv(A).T


def cubic_rotations():
    """The 24 proper rotations of the cube (point group O; crystallography/point_groups.py:123-151): the signed
    permutation matrices of determinant +1."""
    out = []
    for perm in itertools.permutations(range(3)):
        for signs in itertools.product((1, -1), repeat=3):
            m = np.zeros((3, 3))
            for i, j in enumerate(perm):
                m[i, j] = signs[i]
            if np.linalg.det(m) > 0:
                out.append(m)
    return np.array(out)


def hexagonal_rotations():
    """The 12 proper rotations of 622 (point group D6; crystallography/point_groups.py:86-101): rotations by k 60 deg
    about z and 2-fold axes in the xy plane every 30 deg."""
    out = []
    for k in range(6):
        t = np.deg2rad(60.0 * k)
        out.append(np.array([[np.cos(t), -np.sin(t), 0], [np.sin(t), np.cos(t), 0], [0, 0, 1.0]]))
    for k in range(6):
        f = np.deg2rad(30.0 * k)
        out.append(np.array([[np.cos(2 * f), np.sin(2 * f), 0], [np.sin(2 * f), -np.cos(2 * f), 0], [0, 0, -1.0]]))
    return np.array(out)


ROTATIONS = {"cubic": cubic_rotations, "hexagonal": hexagonal_rotations}


def hkl_space_ops(rot, B):
    """The point-group rotations acting on Miller indices, B^-1 R B rounded to integers (material.py:1051-1058 uses
    the equivalent integer tables of point_groups.py:163-268)."""
    Bi = np.linalg.inv(B)
    return [np.rint(Bi @ r @ B).astype(int) for r in rot]


@dataclass
class Reflections:
    hkl: np.ndarray          # (N, 3) family representatives
    multiplicity: np.ndarray
    d_spacing: np.ndarray
    two_theta: np.ndarray    # radians
    h_vecs: np.ndarray       # (N, 3) B @ hkl (1/Angstrom, with 2 pi)


def reflections(A, system, hkl_list, wavelength):
    """Reflection families from an explicit hkl list (material.py:1000-1174): equivalents under the point group and
    Friedel pairs, representative min(equivalents), multiplicity = number of equivalents, sorted by two-theta.
    Pinned by tests/test_oracle_crystal.py (the notebook's Al reflection table: multiplicities 8, 6, 12, 24, 8)."""
    B = reciprocal(A)
    ops = hkl_space_ops(ROTATIONS[system](), B)
    fams, mults, seen = [], [], set()
    for h in hkl_list:
        t = tuple(int(v) for v in h)
        if t in seen:
            continue
        eq = []
        for R in ops:
            w = tuple(int(v) for v in R @ np.asarray(t))
            for cand in (w, tuple(-v for v in w)):
                if cand not in eq:
                    eq.append(cand)
        seen.update(eq)
        fams.append(min(eq))
        mults.append(len(eq))
    hkl = np.array(fams, dtype=np.int64)
    G = (B @ hkl.T).T
    q = np.linalg.norm(G, axis=1)
    tth = 2.0 * np.arcsin(np.clip(q * wavelength / (4.0 * np.pi), 0.0, 1.0))
    order = np.argsort(tth)
    return Reflections(hkl[order], np.array(mults)[order], (2 * np.pi / q)[order], tth[order], G[order])


def group_rings(two_theta, rtol=1e-6):
    """Reflections of equal two-theta form a ring (crystallography/intensity.py:156-169)."""
    rings = []
    for i in np.argsort(two_theta, kind="stable"):
        if rings and np.isclose(two_theta[i], two_theta[rings[-1][0]], rtol=rtol, atol=0.0):
            rings[-1].append(int(i))
        else:
            rings.append([int(i)])
    return rings


def pole_axes(h_vecs, multiplicity, rings, rot):
    """Distinct pole axes of every ring and the number of symmetry images on each
    (single_phase_forward_operator.py:742-826, normalized=True): images S_g h, sign made canonical (first component
    with |.| > 1e-6 positive), merged after rounding to 6 decimals, renormalised.
    Pinned by tests/test_oracle_pf.py (PF from the merged axes == PF summed over all |G| images; counts sum to |G|)."""
    hn = h_vecs / (np.linalg.norm(h_vecs, axis=1, keepdims=True) + 1e-12)
    axes, counts, start = [], [], [0]
    for refl in rings:
        share = multiplicity[refl].astype(np.float64)
        share = share / share.sum()
        imgs, wts = [], []
        for f, w in zip(refl, share):
            img = rot @ hn[f]
            imgs.append(img)
            wts.append(np.full(len(img), w))
        imgs, wts = np.concatenate(imgs), np.concatenate(wts)
        first = np.argmax(np.abs(imgs) > 1e-6, axis=1)
        imgs = imgs * np.sign(imgs[np.arange(len(imgs)), first])[:, None]
        uniq, inv = np.unique(np.round(imgs, 6), axis=0, return_inverse=True)
        cnt = np.bincount(inv.ravel(), weights=wts, minlength=len(uniq))
        axes.append(uniq / np.linalg.norm(uniq, axis=1, keepdims=True))
        counts.append(cnt)
        start.append(start[-1] + len(uniq))
    return np.concatenate(axes), np.concatenate(counts), np.asarray(start)


# ============================================================================ quadrature
def gauss_constant(p):
    """c_p of the p-point Gauss-Legendre error (operators/quadrature.py:32-34). Pinned: c_1 = 1/24, c_2 = 1/4320."""
    return factorial(p) ** 4 / ((2 * p + 1) * factorial(2 * p) ** 3)


def peak_derivative_factor(m, a):
    """mu_m(a), the m-th moment of the symmetric Skellam distribution (quadrature.py:37-48), by the moment-cumulant
    recursion with cumulants a (even) and 0 (odd). Pinned: closed forms mu_2..mu_8 and Bessel sums
    (tests/test_oracle_quadrature.py)."""
    mom = [1.0]
    for n in range(1, m + 1):
        mom.append(sum(comb(n - 1, j - 1) * a * mom[n - j] for j in range(2, n + 1, 2)))
    return mom[m]


def direction_bound(p, k, width, sigma):
    h = abs(float(width)) / k
    return gauss_constant(p) * h ** (2 * p) * peak_derivative_factor(2 * p, 1.0 / float(sigma) ** 2)


def nodes_weights(p, k):
    """k panels of p-point Gauss-Legendre on (0, 1) (quadrature.py:59-63). Pinned: exact for polynomials."""
    g, w = np.polynomial.legendre.leggauss(p)
    x = ((np.arange(k)[:, None] + (g[None, :] + 1.0) / 2.0) / k).ravel()
    return x, np.tile(w / 2.0 / k, k)


def choose_rule(d_omega, d_eta, sigma, tol=0.05, cut_share=0.0, p_max=8, n_max=64):
    """Fewest nodes per bin whose bound meets tol - cut_share (quadrature.py:127-153): returns
    (p_omega, k_omega, p_eta, k_eta). Pinned by the reference test's known answers (test_quadrature.py:67-85)."""
    budget = float(tol) - float(cut_share)
    if budget <= 0:
        raise ValueError("tolerance below the cut share")

    def options(width):
        out = []
        for n in range(1, n_max + 1):
            best = None
            for p in range(1, min(p_max, n) + 1):
                if n % p:
                    continue
                bd = direction_bound(p, n // p, width, sigma)
                if best is None or bd < best[0]:
                    best = (bd, p, n // p)
            out.append((n, *best))
        return out

    best = None
    for nw, bw, pw, kw in options(d_omega):
        for ne, be, pe, ke in options(d_eta):
            if bw + be <= budget:
                key = (nw * ne, bw + be)
                if best is None or key < best[0]:
                    best = (key, (pw, kw, pe, ke))
                break
    if best is None:
        raise ValueError("no quadrature rule meets the tolerance")
    return best[1]


def rule_nodes(a0, a1, n, p, k):
    """Node angles of n bins over [a0, a1] (quadrature.py:204-213): equal sub-bins for p = 1 (the linspace
    expression), Gauss-Legendre panels otherwise; and the per-node weights within a bin."""
    delta = (a1 - a0) / n
    if p == 1:
        nodes = np.linspace(a0, a1, n * k, endpoint=False) + delta / k / 2
        w = np.full(k, 1.0 / k)
    else:
        x, w = nodes_weights(p, k)
        nodes = (a0 + (np.arange(n)[:, None] + x[None, :]) * delta).ravel()
    return nodes, w


# ============================================================================ the operator
def fov_support_mask(Nx, Ny, n_det, angles=None, image_width=None, detector_width=None, detector_shift=0.0):
    """Pixels whose centre projects onto the detector at every angle (utils/support.py:5-51)."""
    image_width = float(Nx if image_width is None else image_width)
    detector_width = float(n_det if detector_width is None else detector_width)
    dx = image_width / max(Nx, Ny)
    x = (np.arange(Nx) - (Nx - 1) / 2.0)[None, :] * dx
    y = (np.arange(Ny) - (Ny - 1) / 2.0)[:, None] * dx
    half = detector_width / 2.0
    if angles is None:
        return x ** 2 + y ** 2 <= (half - abs(float(detector_shift))) ** 2
    mask = np.ones((Ny, Nx), dtype=bool)
    for a in np.asarray(angles, dtype=np.float64):
        t = np.cos(a - 0.5 * np.pi) * x + np.sin(a - 0.5 * np.pi) * y - detector_shift
        mask &= np.abs(t) <= half + 1e-9 * detector_width
    return mask


class OracleOperator:
    """FP64 forward/adjoint of SinglePhaseForwardOperator with the native projector, normalized=True.

    cfg: the reference's cfg dict (Nx, Ny, My, N_Omega, angle_range, cor_offset, N_eta, eta_angle_range, energy,
    k_direction_0, p_direction_0, detector_direction_origin, detector_direction_positive_90).
    lattice: (a, b, c, alpha, beta, gamma) or a 3x3 direct matrix (rows = lattice vectors); system: "cubic" or
    "hexagonal"; hkl_list; wavelength of the material; U (K, 3, 3) orientation matrices (crystal -> sample);
    sigma: kernel width(s) in radians.
    """

    def __init__(self, cfg, lattice, system, hkl_list, wavelength, U, sigma, pf_cutoff_sigma=4.0, quadrature="auto",
                 quadrature_tol=0.05, radon_kernel="trapezoid", orthonormalize=True):
        self.cfg = dict(cfg)
        c = self.cfg
        self.Nx, self.Ny, self.My, self.R = int(c["Nx"]), int(c["Ny"]), int(c["My"]), int(c["N_Omega"])
        self.C = int(c["N_eta"])
        self.cor = float(c.get("cor_offset", 0.0))
        # material (material.py:639-755 with hkl_list)
        if np.ndim(lattice) == 2:
            M = np.asarray(lattice, np.float64)          # rows = direct lattice vectors
            A = M.T
        else:
            A = lattice_matrix(*[float(v) for v in lattice])
        self.refl = reflections(A, system, hkl_list, wavelength)
        self.rings = group_rings(self.refl.two_theta)
        self.P = len(self.rings)
        self.CP = self.C * self.P
        rot = ROTATIONS[system]()
        self.axes, self.counts, self.start = pole_axes(self.refl.h_vecs, self.refl.multiplicity, self.rings, rot)
        # orientations: scipy Rotation round trip as the reference's Grid (utils/grid.py:238-278, 665-700)
        U = np.asarray(U, np.float64)
        self.U = Rotation.from_matrix(U).as_matrix() if orthonormalize else U
        self.K = len(self.U)
        sig = np.broadcast_to(np.asarray(sigma, np.float64), (self.K,))
        self.sigma = sig.astype(np.float32).astype(np.float64)      # the operator stores float32 widths
        self.cut = float(pf_cutoff_sigma) ** 2 / 2 if pf_cutoff_sigma is not None else 6.0
        # angles (single_phase_forward_operator.py:407-414)
        ar = np.array(c["angle_range"], np.float64) / 180 * np.pi
        er = np.array(c.get("eta_angle_range", [0, 360]), np.float64) / 180 * np.pi
        self.angles = np.linspace(ar[0], ar[1], self.R, endpoint=False) + (ar[1] - ar[0]) / self.R / 2
        # quadrature (quadrature.py:168-222)
        d_om, d_eta = (ar[1] - ar[0]) / self.R, (er[1] - er[0]) / self.C
        share = float(np.exp(-self.cut))
        if quadrature == "auto":
            self.rule = choose_rule(d_om, d_eta, float(self.sigma.min()), quadrature_tol, share)
        elif quadrature == "midpoint":
            self.rule = (1, int(c.get("N_Omega_subdivisions", 1)), 1, int(c.get("N_eta_subdivisions", 1)))
        else:
            self.rule = tuple(quadrature)
        pw, kw, pe, ke = self.rule
        self.S, self.Se = pw * kw, pe * ke
        self.om_nodes, ww = rule_nodes(ar[0], ar[1], self.R, pw, kw)
        self.eta_nodes, we = rule_nodes(er[0], er[1], self.C, pe, ke)
        self.qw = ww[:, None] * we[None, :]                          # (S, Se), sums to 1
        # two-theta of the rings: float32 arithmetic on float32 h vectors (single_phase_forward_operator.py:553-557)
        lam = KEV_ANGSTROM / float(c["energy"])
        h32 = np.asarray(self.refl.h_vecs, np.float32)[[r[0] for r in self.rings]]
        self.two_theta = (2.0 * np.arcsin(np.linalg.norm(h32, axis=1) / (4.0 * np.pi) * lam)).astype(np.float32)
        # probed directions v (R, S, C, Se, P, 3) = R(omega)^T u (single_phase_forward_operator.py:569-616),
        # trig of the float32 two-theta in float64 (differentiable/reference.py:131-141)
        k0 = np.asarray(c["k_direction_0"], np.float64)
        p0 = np.asarray(c["p_direction_0"], np.float64)
        d0 = np.asarray(c["detector_direction_origin"], np.float64)
        d90 = np.asarray(c["detector_direction_positive_90"], np.float64)
        eta = self.eta_nodes.reshape(self.C, self.Se)
        e = np.cos(eta)[..., None] * d0 + np.sin(eta)[..., None] * d90                    # (C, Se, 3)
        th = 0.5 * self.two_theta.astype(np.float64)
        u = np.cos(th)[None, None, :, None] * e[:, :, None, :] - np.sin(th)[None, None, :, None] * p0   # (C,Se,P,3)
        Rm = Rotation.from_rotvec(self.om_nodes[:, None] * k0).as_matrix().reshape(self.R, self.S, 3, 3)
        self.v = np.einsum("rsij,cepi->rscepj", Rm, u)
        self.poles = np.einsum("kij,aj->kai", self.U, self.axes)                         # (K, A, 3)
        # projector (operators/parallel_radon.py:92-124)
        self.kernel = radon_kernel
        dx = float(self.Nx) / max(self.Nx, self.Ny)
        rho = 1.0 / dx
        self.X = np.cos(self.angles - 0.5 * np.pi) / rho
        self.Y = np.sin(self.angles - 0.5 * np.pi) / rho
        self.T0 = (self.My - 1) / 2.0 - self.X * (self.Nx - 1) / 2.0 - self.Y * (self.Ny - 1) / 2.0 - self.cor
        self.scale = dx * dx
        self._pf = {}
        self._P = {}
        self.subsets = None

    # ------------------------------------------------------------------ shapes
    @property
    def coeff_shape(self):
        return (self.K, self.Ny, self.Nx)

    @property
    def data_shape(self):
        return (self.R, self.My, self.CP)

    def support_mask(self):
        return fov_support_mask(self.Nx, self.Ny, self.My, self.angles, float(self.Nx), self.My, self.cor)

    # ------------------------------------------------------------------ pole figure
    def pf_angle(self, r, k=None):
        """PF[r] as a dense (K, C, P) float64 array (create_pfo_matrix.py:108-199 / pf_generate.cl:62-140; the
        reference model's pf, differentiable/reference.py:189-211). Only entries near a pole get exp evaluated: the
        cut makes all others exactly zero. Pinned by tests/test_oracle_pf.py (all-|G|-images brute force, sign
        evenness, peak value) and the notebook golden (tests/test_oracle_golden.py)."""
        ks = np.arange(self.K) if k is None else np.asarray(k)
        s = 1.0 / self.sigma[ks] ** 2
        norm = 1.0 / (8.0 * np.pi * self.sigma[ks] ** 2)
        out = np.zeros((len(ks), self.C, self.P))
        for p in range(self.P):
            a0, a1 = self.start[p], self.start[p + 1]
            na = a1 - a0
            V = self.v[r, :, :, :, p, :].reshape(-1, 3)                           # (S*C*Se, 3)
            Pn = self.poles[ks, a0:a1, :].reshape(-1, 3)                          # (K*na, 3)
            ad = np.abs(V @ Pn.T)                                                 # (nodes, K*na)
            sc = np.repeat(s, na)
            # the first term is kept where (1 - |d|) s < cut, i.e. |d| > 1 - cut / s: exp only there
            ii, jj = np.nonzero(ad > (1.0 - self.cut / sc)[None, :])
            x1 = (1.0 - ad[ii, jj]) * sc[jj]
            ok = x1 < self.cut
            g = np.zeros_like(ad)
            g[ii[ok], jj[ok]] = np.exp(-x1[ok])
            if sc.min() < self.cut:                     # the second term (1 + |d|) s < cut: only for very wide kernels
                x2 = (1.0 + ad) * sc[None, :]
                m2 = x2 < self.cut
                g[m2] += np.exp(-x2[m2])
            g = g.reshape(self.S, self.C, self.Se, len(ks), na) * self.counts[a0:a1]
            node = g.sum(axis=-1)                                                 # (S, C, Se, K)
            val = np.einsum("scek,se->kc", node, self.qw)
            out[:, :, p] = val * norm[:, None]
        return out

    def pf_sparse(self, r):
        """PF[r] as a sparse (K, C*P) CSR matrix (cached)."""
        if r not in self._pf:
            self._pf[r] = sp.csr_matrix(self.pf_angle(r).reshape(self.K, self.CP))
        return self._pf[r]

    def prefetch(self, rows=None, workers=None):
        """Evaluate the PF and projection matrices of the given angles (all by default) on a thread pool (NumPy's
        products and exponentials release the GIL); the values are those of pf_sparse / radon_matrix."""
        from concurrent.futures import ThreadPoolExecutor
        rows = [int(r) for r in (range(self.R) if rows is None else rows) if int(r) not in self._pf]
        workers = workers or min(16, len(os.sched_getaffinity(0)))
        with ThreadPoolExecutor(workers) as ex:
            for r, m in zip(rows, ex.map(lambda r: sp.csr_matrix(self.pf_angle(r).reshape(self.K, self.CP)), rows)):
                self._pf[r] = m
        for r in range(self.R) if rows is None else rows:
            self.radon_matrix(r)

    # ------------------------------------------------------------------ projector
    def _kernel(self, d, A):
        if self.kernel == "hat":
            return np.maximum(1.0 - np.abs(d), 0.0)
        hw, top = 0.5 * (1.0 + A), min(1.0, A)
        return np.clip(hw - np.abs(d), 0.0, top) / A

    def radon_matrix(self, r):
        """P[r] (My, Npix) as CSR: S_k[r] = P[r] @ c_k (differentiable/reference.py:225-238; radon_kernels.cl:1-101).
        Every bin a pixel's footprint can reach (|t - s| < (1 + A) / 2 <= 1) is among floor(t) - 1 .. floor(t) + 2.
        Pinned by tests/test_oracle_radon.py (dense brute force over all bins, trapezoid constant row sum, hat ==
        trapezoid at 0/90 degrees)."""
        if r in self._P:
            return self._P[r]
        yy, xx = np.meshgrid(np.arange(self.Ny), np.arange(self.Nx), indexing="ij")
        px, py = xx.ravel().astype(np.float64), yy.ravel().astype(np.float64)
        A = max(abs(self.X[r]), abs(self.Y[r]))
        t = self.X[r] * px + self.Y[r] * py + self.T0[r]
        rows, cols, vals = [], [], []
        base = np.floor(t).astype(np.int64)
        for off in (-1, 0, 1, 2):
            sidx = base + off
            w = self._kernel(t - sidx, A)
            ok = (w > 0) & (sidx >= 0) & (sidx < self.My)
            rows.append(sidx[ok])
            cols.append(np.nonzero(ok)[0])
            vals.append(self.scale * w[ok])
        M = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                          shape=(self.My, self.Nx * self.Ny))
        self._P[r] = M
        return M

    # ------------------------------------------------------------------ forward / adjoint
    def _angles(self, subset):
        return np.arange(self.R) if subset is None else self.subsets[subset]

    def direct(self, x, subset=None):
        """A x (R, My, C P); with subset m only its rows (others 0) (single_phase_forward_operator.py:1646-1729)."""
        d = np.zeros(self.data_shape)
        rows = self._angles(subset)
        d[rows] = self.direct_rows(x, rows)
        return d

    def _stacked(self, rows):
        """The projection matrices of the given angles stacked, (len(rows) My, Npix), and its transpose (cached)."""
        key = tuple(int(r) for r in rows)
        c = self.__dict__.setdefault("_stack", {})
        if key not in c:
            P = sp.vstack([self.radon_matrix(r) for r in key]).tocsr()
            c[key] = (P, P.T.tocsr())
        return c[key]

    def direct_rows(self, x, rows):
        """(len(rows), My, C P): the rows of A x of the given angles."""
        xt = np.ascontiguousarray(np.asarray(x, np.float64).reshape(self.K, -1).T)     # (Npix, K)
        S = (self._stacked(rows)[0] @ xt).reshape(len(rows), self.My, self.K)
        out = np.empty((len(rows), self.My, self.CP))
        for i, r in enumerate(rows):
            out[i] = (self.pf_sparse(r).T @ S[i].T).T        # (My, CP)
        return out

    def adjoint(self, y, subset=None):
        """A^T y (K, Ny, Nx), the exact transpose (single_phase_forward_operator.py:1800-2033: alpha R/pi times the
        backprojection weight pi/R is 1); with subset m only its rows are read."""
        yf = np.asarray(y, np.float64)
        rows = self._angles(subset)
        Z = np.empty((len(rows), self.My, self.K))
        for i, r in enumerate(rows):
            Z[i] = (self.pf_sparse(r) @ yf[r].T).T           # (My, K)
        out = self._stacked(rows)[1] @ Z.reshape(-1, self.K)  # (Npix, K)
        return np.ascontiguousarray(out.T).reshape(self.coeff_shape)

    def set_omega_subsets(self, subsets):
        """Interleaved subsets np.arange(m, R, M) or given lists, each sorted and unique
        (single_phase_forward_operator.py:1541-1599)."""
        if np.isscalar(subsets):
            M = int(subsets)
            subsets = [np.arange(m, self.R, M) for m in range(M)]
        self.subsets = [np.unique(np.asarray(s, dtype=np.int64)) for s in subsets]
        return self.subsets


# ============================================================================ Poisson pieces
def periodic(a, n):
    """A per-segment / per-(translation, segment) / per-point array expanded to n points (fista_poisson.py:287-308)."""
    if a is None:
        return None
    a = np.asarray(a, np.float64).ravel()
    return a[np.arange(n) % a.size]


def deviance(mu, b, w=None, g=None, bg=None, eps=1e-3):
    """Poisson deviance sum w/g (m - b + b log(b/m)), m = max(mu, 0) + bg + eps g, over points with w != 0
    (fista_poisson.py:72-109, mode 0). Pinned by the notebook golden objectives (tests/test_oracle_golden.py)."""
    mu = np.asarray(mu, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    n = mu.size
    w = np.ones(n) if w is None else periodic(w, n)
    g = np.ones(n) if g is None else periodic(g, n)
    bg = np.zeros(n) if bg is None else periodic(bg, n)
    eps32 = float(np.float32(eps))
    m = np.maximum(mu, 0.0) + bg + eps32 * g
    v = m - b
    pos = b > 0
    v[pos] += b[pos] * np.log(b[pos] / m[pos])
    cnt = w != 0
    return float(np.sum((w[cnt] / g[cnt]) * v[cnt]))


def mean_counted(b, w=None):
    """Mean of b over points with w != 0 (pdhg_poisson.py:97-111)."""
    b = np.asarray(b, np.float64).ravel()
    w = np.ones(b.size) if w is None else periodic(w, b.size)
    c = w != 0
    return float(b[c].sum() / max(c.sum(), 1))


def dual_prox(y, Ax, s, b, w, g, bg, eps):
    """y' = prox_{s F*}(y + s Ax), F_i(u) = sc (u + r - b log(u + r)), sc = w/g, r = bg + eps g; 0 where w = 0 or
    s = 0 (spdhg_poisson.py:69-91; pdhg_poisson.py:33-46). Pinned: tests/test_oracle_spdhg.py (the prox optimality
    condition, and the solver reaching an independent minimiser)."""
    sc = w / g
    r = bg + eps * g
    v = y + s * Ax + s * r
    dd = v - sc
    yn = 0.5 * (v + sc - np.sqrt(dd * dd + 4.0 * s * sc * b))
    return np.where((w != 0) & (s != 0), yn, 0.0)


TV_SAFETY = 0.99          # spdhg_poisson.py:277: the TV dual step is this fraction of its bound


def _tv_free(T):
    """Free forward differences (right, down) of (K, Ny, Nx) images: both ends have T > 0 (spdhg_poisson.py:110-113)."""
    vr = np.zeros(T.shape, bool)
    vd = np.zeros(T.shape, bool)
    vr[:, :, :-1] = (T[:, :, :-1] > 0) & (T[:, :, 1:] > 0)
    vd[:, :-1, :] = (T[:, :-1, :] > 0) & (T[:, 1:, :] > 0)
    return vr, vd


def _fwd_diff(x):
    gx = np.zeros_like(x)
    gy = np.zeros_like(x)
    gx[:, :, :-1] = x[:, :, 1:] - x[:, :, :-1]
    gy[:, :-1, :] = x[:, 1:, :] - x[:, :-1, :]
    return gx, gy


def tv_dual(x, T, qx, qy, lam, ps):
    """The TV block's dual step (spdhg_poisson.py:110-129 tv_dual_core, 146-165 tv_dual_dense): per pixel
    s = ps / (4 max over its free differences of (T_j + T_j')), q' = proj_{|q| <= lam}(q + s D x) on the free
    differences (0 on the others, and where the pixel has none). Pinned by test_oracle_tv.py (an independent
    Chambolle-Pock on the explicit matrix reaches the same objective)."""
    vr, vd = _tv_free(T)
    Tr = np.zeros_like(T)
    Td = np.zeros_like(T)
    Tr[:, :, :-1] = T[:, :, 1:]
    Td[:, :-1, :] = T[:, 1:, :]
    m = np.zeros_like(T)
    m = np.where(vr, np.maximum(m, T + Tr), m)
    m = np.where(vd, np.maximum(m, T + Td), m)
    gx, gy = _fwd_diff(x)
    s = np.where(m > 0, ps / (4.0 * np.where(m > 0, m, 1.0)), 0.0)
    ax = np.where((m > 0) & vr, qx + s * gx, 0.0)
    ay = np.where((m > 0) & vd, qy + s * gy, 0.0)
    nrm = np.sqrt(ax * ax + ay * ay)
    c = np.where(nrm > lam, lam / np.where(nrm > 0, nrm, 1.0), 1.0)
    return ax * c, ay * c


def tv_adjoint(dqx, dqy):
    """D^T dq (spdhg_poisson.py:168-179 tv_adjoint_dense): -dqx - dqy + dqx(left) + dqy(up)."""
    v = -dqx - dqy
    v[:, :, 1:] += dqx[:, :, :-1]
    v[:, 1:, :] += dqy[:, :-1, :]
    return v


def tv_value(x, T):
    """sum_k TV(x_k) over the free differences (spdhg_poisson.py:131-137, 181-193)."""
    vr, vd = _tv_free(T)
    gx, gy = _fwd_diff(x)
    return float(np.sqrt(np.where(vr, gx, 0.0) ** 2 + np.where(vd, gy, 0.0) ** 2).sum())


# ============================================================================ differentiable model pieces
def intensity_forward(u, angle_scale, trans_scale, ring_scale, background, n_rings):
    """The intensity layer d[a, s, j] = f_a t_s r_{p(j)} u[a, s, j] + b_j for data (R, My, C P), segment j = c P + p,
    so p(j) = j mod P (differentiable/data_model.py:40-43; the segment order of utils/arrays.py:3-8). Written over
    the flat segment axis with the ring scale tiled C times. Pinned by tests/test_oracle_intensity.py: a one-hot
    angle / translation / ring scale touches exactly its rows / points / segments, the sum over unit inputs has the
    closed form (sum f)(sum t)(C sum r) + R My sum b, and the vjp equals the exact directional derivatives."""
    u = np.asarray(u, np.float64)
    R, My, CP = u.shape
    C = CP // int(n_rings)
    rs = np.tile(np.asarray(ring_scale, np.float64), C)                   # r_{j mod P}
    f = np.asarray(angle_scale, np.float64)
    t = np.asarray(trans_scale, np.float64)
    return f[:, None, None] * t[None, :, None] * rs[None, None, :] * u + np.asarray(background, np.float64)


def intensity_vjp(u, angle_scale, trans_scale, ring_scale, r, n_rings):
    """Gradients of <r, d> of intensity_forward (data_model.py:44-64): d/df_a = sum_{s,j} r t_s r_p u,
    d/dt_s = sum_{a,j} r f_a r_p u, d/dr_p = sum over a, s and the segments of ring p of r f_a t_s u,
    d/db_j = sum_{a,s} r, and d/du = r f_a t_s r_p."""
    u = np.asarray(u, np.float64)
    r = np.asarray(r, np.float64)
    R, My, CP = u.shape
    P = int(n_rings)
    C = CP // P
    f = np.asarray(angle_scale, np.float64)
    t = np.asarray(trans_scale, np.float64)
    rp = np.asarray(ring_scale, np.float64)
    rs = np.tile(rp, C)
    ru = r * u
    g_f = np.array([np.sum(ru[a] * t[:, None] * rs[None, :]) for a in range(R)])
    g_t = np.array([np.sum(ru[:, s] * f[:, None] * rs[None, :]) for s in range(My)])
    g_r = np.array([np.sum(ru[:, :, p::P] * f[:, None, None] * t[None, :, None]) for p in range(P)])
    g_b = r.sum(axis=(0, 1))
    g_u = r * f[:, None, None] * t[None, :, None] * rs[None, None, :]
    return {"angle_scale": g_f, "trans_scale": g_t, "ring_scale": g_r, "background": g_b, "u": g_u}


def smooth_tv(c, eps=1e-8):
    """Smoothed isotropic TV of every orientation image of c (K, Ny, Nx): sum over pixels of
    sqrt(dx^2 + dy^2 + eps) with forward differences, zero beyond the last row / column (the C4 loss term; the
    reference's TV, spdhg_poisson.py:131-137, without the free-difference mask and smoothed so that it has a
    gradient everywhere). Pinned by tests/test_oracle_intensity.py (closed forms on ramps and constants)."""
    c = np.asarray(c, np.float64)
    dx = np.zeros_like(c)
    dy = np.zeros_like(c)
    dx[:, :, :-1] = np.diff(c, axis=2)
    dy[:, :-1, :] = np.diff(c, axis=1)
    return float(np.sqrt(dx * dx + dy * dy + eps).sum())


class OracleSPDHG:
    """SPDHGPoisson (dense layout, every array resident) in float64 (spdhg_poisson.py:298-751), with the total
    variation block (tv > 0).

    The control flow, the random subset order (numpy default_rng(seed).permutation per epoch, the first subset of an
    epoch never the last of the previous one; with tv the TV block inserted at rng.choice places), the step sizes
    and the block probabilities follow the reference; the arithmetic is float64."""

    def __init__(self, op, subsets=8, lam=0.0, prox_kind="nonneg", support="fov", eps_counts=1e-3, rho="auto", seed=0,
                 tv=0.0, tv_prob=None):
        self.op = op
        self.subsets = op.set_omega_subsets(subsets)
        self.M = len(self.subsets)
        # spdhg_poisson.py:339-346: block probabilities
        self.tv = float(tv)
        self.tv_prob = (1.0 / (self.M + 1) if tv_prob is None else float(tv_prob)) if self.tv > 0 else 0.0
        self._n_tv = max(1, int(round(self.M * self.tv_prob / (1 - self.tv_prob)))) if self.tv > 0 else 0
        self._p_data = (1.0 - self.tv_prob) / self.M
        self._inv_p_data = float(self.M) if self.tv == 0 else 1.0 / self._p_data
        self.lam = float(lam) if prox_kind == "nonneg_l1" else 0.0
        self.eps = float(np.float32(eps_counts))
        self.rho = rho
        self.rng = np.random.default_rng(seed)
        self._last = None
        if support is None or support is False:
            self.supp = None
        elif isinstance(support, str):
            self.supp = op.support_mask()
        else:
            self.supp = np.asarray(support, bool)
        self.state = None

    def _epoch_order(self):
        """spdhg_poisson.py:363-370."""
        perm = self.rng.permutation(self.M)
        if self.M > 1 and int(perm[0]) == self._last:
            perm[[0, 1]] = perm[[1, 0]]
        self._last = int(perm[-1])
        return [int(m) for m in perm]

    def _epoch_sequence(self):
        """spdhg_poisson.py:372-381: with tv, the TV block (-1) at _n_tv random places."""
        order = self._epoch_order()
        if self.tv <= 0:
            return order
        n = len(order) + self._n_tv
        tv_at = set(int(i) for i in self.rng.choice(n, self._n_tv, replace=False))
        it = iter(order)
        return [-1 if i in tv_at else next(it) for i in range(n)]

    def run(self, x0, b, n_epochs, weights=None, gain=None, background=None, diagnostics_interval=1, state=None,
            record=False):
        op = self.op
        n = int(np.prod(op.data_shape))
        b = np.asarray(b, np.float64).reshape(op.data_shape)
        w = np.ones(n) if weights is None else periodic(weights, n)
        g = np.ones(n) if gain is None else periodic(gain, n)
        bg = np.zeros(n) if background is None else periodic(background, n)
        w, g, bg = (a.reshape(op.data_shape) for a in (w, g, bg))
        # rho (spdhg_poisson.py:558-567; RHO_AUTO = 3e4, pdhg_poisson.py:33)
        if self.rho == "auto":
            mb = mean_counted(b, w)
            rho = 3e4 / mb ** 2 if mb > 0 else 1.0
        else:
            rho = float(self.rho)
        self.rho_used = rho
        gamma = np.sqrt(rho)
        p = self._p_data
        # the support: (Ny, Nx) for every orientation, or (K, Ny, Nx) per coefficient (a local basis)
        sup = np.ones((op.Ny, op.Nx), bool) if self.supp is None else self.supp
        supb = np.broadcast_to(sup, op.coeff_shape)
        # S from the row sums of A 1_S (spdhg_poisson.py:621-631; dual_steps pdhg_poisson.py:79-84)
        ones = supb.astype(np.float64)
        D = op.direct(ones)
        sig = np.where((w != 0) & (D > 0), gamma / np.where(D > 0, D, 1.0), 0.0)
        # T = min_m p / (gamma A_m^T 1_counted) on the support (spdhg_poisson.py:632-643, 252-272)
        mask = (w != 0).astype(np.float64)
        T = np.full(op.coeff_shape, np.inf)
        for m in range(self.M):
            col = op.adjoint(mask, subset=m)
            ok = (col > 0) & supb
            T = np.where(ok, np.minimum(T, (p / gamma) / np.where(col > 0, col, 1.0)), T)
        T[np.isinf(T)] = 0.0
        self.T, self.sig = T, sig
        x = np.maximum(np.asarray(x0, np.float64).reshape(op.coeff_shape), 0.0) * supb
        y = np.zeros(op.data_shape) if state is None else np.asarray(state["y"], np.float64).reshape(op.data_shape)
        if state is not None and state.get("z") is not None:
            z = np.asarray(state["z"], np.float64).reshape(op.coeff_shape).copy()
        elif state is not None:
            z = op.adjoint(y)
        else:
            z = np.zeros(op.coeff_shape)
        ep0 = int(state["epoch"]) if state is not None else 0
        inv_p = self._inv_p_data
        if self.tv > 0:
            qx = np.zeros(op.coeff_shape) if state is None or state.get("qx") is None else \
                np.asarray(state["qx"], np.float64).reshape(op.coeff_shape).copy()
            qy = np.zeros(op.coeff_shape) if state is None or state.get("qy") is None else \
                np.asarray(state["qy"], np.float64).reshape(op.coeff_shape).copy()
        self.iter_stats = []
        self.iterates = []
        f = float("nan")

        def dual(m, D):
            rows = self.subsets[m]
            yn = dual_prox(y[rows], D[rows], sig[rows], b[rows], w[rows], g[rows], bg[rows], self.eps)
            D[rows] = yn - y[rows]
            y[rows] = yn

        def update(dz, ip):
            nonlocal x, z
            z = z + dz
            zb = z + ip * dz
            x = np.where(T > 0, np.maximum(0.0, x - T * (zb + self.lam)), 0.0)

        # spdhg_poisson.py:664-697
        seq = self._epoch_sequence()
        D = np.zeros(op.data_shape)
        if seq[0] >= 0:
            D[self.subsets[seq[0]]] = op.direct_rows(x, self.subsets[seq[0]])
        for e in range(n_epochs):
            last = e + 1 == n_epochs
            want_f = (e + 1) % max(1, diagnostics_interval) == 0 or last
            nseq = None
            for j, blk in enumerate(seq):
                end = j + 1 == len(seq)
                if not end:
                    nb = seq[j + 1]
                elif not last:
                    nseq = self._epoch_sequence()
                    nb = nseq[0]
                else:
                    nb = None
                adj = blk if blk >= 0 else None
                m_next = nb if nb is not None and nb >= 0 else None
                if adj is not None:
                    dual(adj, D)
                    update(op.adjoint(D, subset=adj), inv_p)
                else:
                    nqx, nqy = tv_dual(x, T, qx, qy, self.tv, TV_SAFETY * self.tv_prob)
                    dz = tv_adjoint(nqx - qx, nqy - qy)
                    qx, qy = nqx, nqy
                    update(dz, 1.0 / self.tv_prob)
                if record:
                    self.iterates.append(x.copy())
                if (end and want_f) or nb is None or (m_next is not None and m_next == adj):
                    if (end and want_f) or m_next is not None:
                        D = op.direct(x)
                        if end and want_f:
                            f = deviance(D, b, w, g, bg, self.eps)
                            if self.tv > 0:
                                f += self.tv * tv_value(x, T)
                elif m_next is not None:
                    rows = self.subsets[m_next]
                    D[rows] = op.direct_rows(x, rows)
            if nseq is not None:
                seq = nseq
            self.iter_stats.append({"epoch": ep0 + e + 1, "f": f if want_f else None, "rho": rho})
        self.state = {"epoch": ep0 + n_epochs, "z": z, "y": y}
        if self.tv > 0:
            self.state["qx"], self.state["qy"] = qx, qy
        self.final_stats = {"epochs": n_epochs, "final_f": f, "rho": rho, "subsets": self.M, "tv": self.tv,
                            "tv_prob": self.tv_prob}
        return x


def estimate_L_power(op, niter=20, seed=0, eps=1e-30):
    """Power iteration for ||A^T A|| (single_phase_forward_operator.py:2136-2206): start drawn as (Nx, Ny, K) float32
    normals and transposed; float64 dot products here. Pinned by tests/test_oracle_power.py (the largest eigenvalue
    of the explicit A^T A on a tiny operator)."""
    rng = np.random.default_rng(seed)
    x = np.ascontiguousarray(rng.standard_normal((op.Nx, op.Ny, op.K)).astype(np.float32).transpose(2, 1, 0))
    x = x.astype(np.float64)
    x /= np.sqrt(np.vdot(x, x)) + eps
    L = 0.0
    for _ in range(niter):
        z = op.adjoint(op.direct(x))
        L = float(np.vdot(x, z) / (np.vdot(x, x) + eps))
        z /= np.sqrt(np.vdot(z, z)) + eps
        x = z
    return L


# ============================================================================ OSEM
class OracleOSEM:
    """Ordered-subsets EM / MLEM in float64 (optimization/osem_poisson.py:40-185; poisson_term.py:39-47 for the EM
    ratio w b / (g mu) and the sensitivity weights w / g): the same start rule (start_floor x the median positive
    allowed coefficient, or the flat start matching the fitted data's total), the same subset order
    (default_rng(seed).permutation(M) per epoch), angles in no subset unfitted (weight 0).
    Pinned by tests/test_oracle_osem.py: MLEM's multiplicative update preserves the fitted total prediction
    (sum_j sens_j x_j' = sum_i b_i (A x)_i / mu_i), never increases the deviance (the EM ascent property), keeps
    x >= 0 and 0 off the support, and held-out angles' data leave the fit unchanged."""

    def __init__(self, op, subsets=8, lam=0.0, support="fov", eps_counts=1e-3, seed=0):
        self.op = op
        if subsets is None:
            self.subsets = None
            self.M = 1
        else:
            self.subsets = op.set_omega_subsets(subsets)
            self.M = len(self.subsets)
        self.lam = float(lam)
        self.eps = float(np.float32(eps_counts))
        self.rng = np.random.default_rng(seed)
        if support is None:
            self.supp = None
        elif isinstance(support, str):
            self.supp = np.asarray(op.support_mask(), bool)
        else:
            self.supp = np.asarray(support) > 0
        self.iter_stats = []
        self.final_f = None
        self.start_floor_used = None

    def _sub(self, m):
        return None if self.subsets is None else m

    def run(self, x0, b, n_epochs, weights=None, gain=None, background=None, start_floor=0.01,
            diagnostics_interval=1, record=False):
        op = self.op
        dshape, cshape = tuple(op.data_shape), tuple(op.coeff_shape)
        n = int(np.prod(dshape))
        b = np.asarray(b, np.float64).reshape(dshape)
        w = np.ones(n) if weights is None else periodic(weights, n)
        g = np.ones(n) if gain is None else periodic(gain, n)
        bg = np.zeros(n) if background is None else periodic(background, n)
        w, g, bg = (a.reshape(dshape) for a in (w, g, bg))
        if self.subsets is not None:
            fit = np.zeros(dshape[0], bool)
            for s in self.subsets:
                fit[s] = True
            w = w * fit[:, None, None]
        wg = np.where(w != 0, w / g, 0.0)
        sens = [op.adjoint(wg, subset=self._sub(m)) for m in range(self.M)]
        allowed = np.ones(cshape, bool) if self.supp is None else np.broadcast_to(self.supp, cshape)
        x = np.where(allowed, np.asarray(x0, np.float64).reshape(cshape), 0.0)
        if start_floor:
            pos = x[x > 0]
            if pos.size:
                floor = float(start_floor) * float(np.median(pos))
            else:
                st = sum(float(s[allowed].sum()) for s in sens)
                tot = float((wg * b).sum()) - float((wg * bg).sum())
                floor = max(tot, 0.0) / st if st > 0 else 1.0
            x = np.where(allowed, np.maximum(x, floor), 0.0)
            self.start_floor_used = floor
        self.iter_stats = []
        self.iterates = []
        f = None

        def dev_full(xv):
            return deviance(op.direct(xv), b, w, g, bg, self.eps)

        for ep in range(int(n_epochs)):
            order = self.rng.permutation(self.M) if self.M > 1 else [0]
            for m in order:
                m = int(m)
                pred = op.direct(x, subset=self._sub(m))
                mu = np.maximum(pred, 0.0) + bg + self.eps * g
                r = np.where(w != 0, wg * b / mu, 0.0)
                back = op.adjoint(r, subset=self._sub(m))
                s = sens[m]
                x = np.where(~allowed, 0.0, np.where(s > 0, x * back / np.where(s > 0, s + self.lam, 1.0), x))
                if record:
                    self.iterates.append(x.copy())
            stats = {"epoch": ep + 1}
            f = None
            if diagnostics_interval and ((ep + 1) % int(diagnostics_interval) == 0 or ep + 1 == int(n_epochs)):
                f = dev_full(x)
                stats["f"] = f
            self.iter_stats.append(stats)
        if f is None:
            f = dev_full(x)
        self.final_f = f
        return x


# ============================================================================ PDHG
class OraclePDHG:
    """Chambolle-Pock with diagonal preconditioning in float64 (optimization/pdhg_poisson.py:120-265, the GPU-resident
    path): Sigma = sigma0 / (A 1_S) on counted points, T = tau0 / (A^T 1_counted) on the support, tau0 sigma0 = 1,
    rho = sigma0 / tau0 (auto: 3e4 / mean(b)^2 over counted points); y <- prox_{Sigma F*}(y + Sigma A x_bar) (the
    dual_prox above); x' = max(0, x - T (A^T y + lam)) where T > 0, else 0; x_bar = 2 x' - x. The objective is the
    deviance of A x of the iterate (exact here: no float32 recursion). Pinned by tests/test_oracle_pdhg.py: the
    L-BFGS-B minimiser reached (the reference's own bar, tests/test_poisson.py:189-207: the best of three step
    balances within 1e-3 after 4000 iterations), the nonneg_l1 minimiser's KKT conditions, and a resumed run equal
    to one run."""

    def __init__(self, op, prox_kind="nonneg", lam=0.0, support="fov", eps_counts=1e-3, rho="auto"):
        self.op = op
        self.lam = float(lam) if prox_kind == "nonneg_l1" else 0.0
        self.eps = float(np.float32(eps_counts))
        self.rho = rho
        if support is None:
            self.supp = None
        elif isinstance(support, str):
            self.supp = np.asarray(op.support_mask(), bool)
        else:
            self.supp = np.asarray(support) > 0
        self.iter_stats = []

    def run(self, x0, b, niter, weights=None, gain=None, background=None, state=None):
        op = self.op
        dshape, cshape = tuple(op.data_shape), tuple(op.coeff_shape)
        n = int(np.prod(dshape))
        b = np.asarray(b, np.float64).reshape(dshape)
        w = np.ones(n) if weights is None else periodic(weights, n)
        g = np.ones(n) if gain is None else periodic(gain, n)
        bg = np.zeros(n) if background is None else periodic(background, n)
        w, g, bg = (a.reshape(dshape) for a in (w, g, bg))
        if self.rho == "auto":
            mb = mean_counted(b, w)
            rho = 3e4 / mb ** 2 if mb > 0 else 1.0
        else:
            rho = float(self.rho)
        self.rho_used = rho
        tau0, sigma0 = 1.0 / np.sqrt(rho), np.sqrt(rho)
        sup = np.ones((op.Ny, op.Nx), bool) if self.supp is None else self.supp
        rows = op.direct(np.broadcast_to(sup, cshape).astype(np.float64))
        sig = np.where((w != 0) & (rows > 0), sigma0 / np.where(rows > 0, rows, 1.0), 0.0)
        cols = op.adjoint((w != 0).astype(np.float64))
        ok = (cols > 0) & sup[None]
        T = np.where(ok, tau0 / np.where(cols > 0, cols, 1.0), 0.0)
        x = np.maximum(np.asarray(x0, np.float64).reshape(cshape), 0.0) * sup[None]
        if state is not None:
            xb = np.asarray(state["x_bar"], np.float64).reshape(cshape).copy()
            y = np.asarray(state["y"], np.float64).reshape(dshape).copy()
        else:
            xb, y = x.copy(), np.zeros(dshape)
        it0 = int(state["it"]) if state is not None else 0
        self.iter_stats = []
        f = deviance(op.direct(x), b, w, g, bg, self.eps)
        for k in range(int(niter)):
            y = dual_prox(y, op.direct(xb), sig, b, w, g, bg, self.eps)
            grad = op.adjoint(y)
            xn = np.where(T > 0, np.maximum(0.0, x - T * (grad + self.lam)), 0.0)
            xb = 2.0 * xn - x
            x = xn
            f = deviance(op.direct(x), b, w, g, bg, self.eps)
            self.iter_stats.append({"iter": it0 + k + 1, "f": f})
        self.state = {"it": it0 + int(niter), "x_bar": xb, "y": y}
        self.final_f = f
        return x


# ============================================================================ FISTA
FISTA_REFRESH = (5, 10, 20, 40, 80, 160, 320, 640, 1280, 2560, 5120)      # fista_poisson.py:284


class OracleFISTA:
    """FISTA in the SQS metric with backtracking, in float64 (optimization/fista_poisson.py:311-617, the
    GPU-resident path; its kernels 61-282): the weighted-least-squares start phase (mode 1: the quadratic
    approximation of the deviance at the data, weights w / (g max(b, floor g))), the metric D = 1 / A^T(c A 1_S),
    the descent test, the metric at y on the first rejection then halving, function-value restarts, the refresh
    schedule, the adjoint-scale estimate. Pinned by tests/test_oracle_fista.py: the L-BFGS-B minimiser reached (the
    reference's bar for its Poisson solvers), the descent of every accepted step, a continued run equal to one run."""

    def __init__(self, op, prox_kind="nonneg", lam=0.0, support="fov", eps_counts=1e-3, floor_counts=1.0,
                 refresh=FISTA_REFRESH, refresh_on_failure=True, max_backtracks=60, start="wls", n_start=20):
        self.op = op
        self.lam = float(lam) if prox_kind == "nonneg_l1" else 0.0
        self.eps = float(np.float32(eps_counts))
        self.floor = float(np.float32(floor_counts))
        self.refresh = tuple(int(r) for r in refresh)
        self.refresh_on_failure = bool(refresh_on_failure)
        self.max_backtracks = int(max_backtracks)
        self.start = start
        self.n_start = int(n_start)
        if support is None:
            self.supp = None
        elif isinstance(support, str):
            self.supp = np.asarray(op.support_mask(), bool)
        else:
            self.supp = np.asarray(support) > 0
        self.adjoint_scale = 1.0

    def _value_grad(self, mu, mode):
        b, w, g, bg = self.b, self.w, self.g, self.bg
        cnt = w != 0
        sc = np.where(cnt, w / g, 0.0)
        if mode == 0:
            m = np.maximum(mu, 0.0) + bg + self.eps * g
            v = m - b + np.where(b > 0, b * np.log(np.where(b > 0, b, 1.0) / m), 0.0)
            return float(np.sum(sc * v)), np.where(cnt, sc * (1.0 - b / m), 0.0)
        r = mu + bg - b
        c = sc / np.maximum(b, self.floor * g)
        return float(np.sum(0.5 * c * r * r)), np.where(cnt, c * r, 0.0)

    def _metric(self, mu, mode):
        op, b, w, g, bg = self.op, self.b, self.w, self.g, self.bg
        cnt = w != 0
        ones = np.broadcast_to(self.sup, op.coeff_shape).astype(np.float64)
        r = op.direct(ones)
        if mode == 0:
            m = np.maximum(np.maximum(mu, 0.0) + bg, self.floor * g)
            c = np.where(cnt, w / g * np.maximum(b, m) / (m * m), 0.0)
        else:
            c = np.where(cnt, w / (g * np.maximum(b, self.floor * g)), 0.0)
        u_r = float(np.sum(c * r * r))
        atu = op.adjoint(c * r)
        at_u = float(atu[np.broadcast_to(self.sup, op.coeff_shape)].sum())
        if u_r > 0 and at_u > 0:
            self.adjoint_scale = at_u / u_r
        return np.where(np.broadcast_to(self.sup, op.coeff_shape) & (atu > 0), 1.0 / np.where(atu > 0, atu, 1.0), 0.0)

    def run(self, x0, b, niter, weights=None, gain=None, background=None, state=None):
        op = self.op
        dshape, cshape = tuple(op.data_shape), tuple(op.coeff_shape)
        n = int(np.prod(dshape))
        self.b = np.asarray(b, np.float64).reshape(dshape)
        self.w = (np.ones(n) if weights is None else periodic(weights, n)).reshape(dshape)
        self.g = (np.ones(n) if gain is None else periodic(gain, n)).reshape(dshape)
        self.bg = (np.zeros(n) if background is None else periodic(background, n)).reshape(dshape)
        self.sup = np.ones((op.Ny, op.Nx), bool) if self.supp is None else self.supp
        x = np.where(np.broadcast_to(self.sup, cshape), np.maximum(np.asarray(x0, np.float64).reshape(cshape), 0.0),
                     0.0)
        zero_start = state is None and x.max() == 0.0
        wls_until = 0
        if self.start == "wls" and self.n_start > 0:
            if zero_start:
                wls_until = self.n_start
            elif state is not None and state.get("wls_until", 0) > int(state["it"]):
                wls_until = int(state["wls_until"])
        mode = 1 if wls_until else 0
        if zero_start and self.start == "uniform":
            ones = np.broadcast_to(self.sup, cshape).astype(np.float64)
            r = op.direct(ones)
            sb, sa = float((self.w * self.b).sum()), float((self.w * r).sum())
            if sa > 0 and sb > 0:
                x = ones * (sb / sa)
        it0, t_mom, theta = 0, 1.0, 1.0
        if state is not None:
            it0, t_mom, theta = int(state["it"]), float(state["t"]), float(state["theta"])
            y = np.asarray(state["y"], np.float64).reshape(cshape).copy()
        else:
            y = x.copy()
        Ax = op.direct(x)
        Ay = op.direct(y)
        D = self._metric(Ax, mode)
        f_x = self._value_grad(Ax, mode)[0]
        F_x = f_x + self.lam * x.sum()
        self.iter_stats = []
        for k in range(int(niter)):
            it = it0 + k + 1
            f_y, gmu = self._value_grad(Ay, mode)
            grad = op.adjoint(gmu)
            theta = min(1.0, 2.0 * theta)
            n_bt, local = 0, False
            while True:
                xn = np.where(D > 0, np.maximum(0.0, y - theta * D * (grad + self.lam)), 0.0)
                Axn = op.direct(xn)
                f_n = self._value_grad(Axn, mode)[0]
                d = xn - y
                lin = float(np.sum(grad * d)) / self.adjoint_scale
                quad = float(np.sum(np.where(D > 0, d * d / np.where(D > 0, D, 1.0), 0.0))) / (theta * self.adjoint_scale)
                if f_n <= f_y + lin + 0.5 * quad + 1e-6 * abs(f_y) or quad == 0.0:
                    break
                n_bt += 1
                if not local and self.refresh_on_failure:
                    D = self._metric(op.direct(y), mode)
                    local, theta = True, 1.0
                    continue
                theta *= 0.5
                if n_bt > self.max_backtracks:
                    raise RuntimeError("no sufficient decrease")
            F_n = f_n + self.lam * xn.sum()
            restarted = F_n > F_x
            if restarted:
                t_mom = 1.0
            t_new = 0.5 * (1.0 + np.sqrt(1.0 + 4.0 * t_mom * t_mom))
            beta = (t_mom - 1.0) / t_new
            t_mom = t_new
            y = xn + beta * (xn - x)
            Ay = Axn + beta * (Axn - Ax)
            x, Ax = xn, Axn
            f_x, F_x = f_n, F_n
            switch = mode == 1 and it >= wls_until
            if switch:
                mode = 0
                f_x = self._value_grad(Ax, 0)[0]
                F_x = f_x + self.lam * x.sum()
            refreshed = (it in self.refresh or switch) and k + 1 < niter and mode == 0
            if refreshed:
                D = self._metric(Ax, 0)
                t_mom, theta = 1.0, 1.0
                y, Ay = x.copy(), Ax.copy()
            self.iter_stats.append({"iter": it, "f": f_n, "obj": F_n, "backtracks": n_bt, "restart": bool(restarted),
                                    "metric_refresh": bool(refreshed), "phase": "wls" if it <= wls_until else "poisson"})
        self.state = {"it": it0 + int(niter), "t": t_mom, "theta": theta, "y": y, "wls_until": wls_until}
        self.final_f, self.final_obj = f_x, F_x
        return x
