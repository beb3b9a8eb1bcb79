"""Material: lattice, reflection families and the proper point group, as the single-phase operator needs them
(the reference's crystallography/material.py, point_groups.py, intensity.py; only what the configured path uses:
no atomic basis, CIF or structure factors, so the operator runs with normalized=True)."""
from __future__ import annotations

import itertools

import numpy as np

_REFLECTION_DTYPE = np.dtype([("hkl", "i4", (3,)), ("multiplicity", "i4"), ("d_spacing", "f8"),
                              ("two_theta", "f8"), ("intensity", "f8"), ("sf_squared", "f8")])


def _cube_group():
    """O: the 24 signed permutation matrices with determinant +1 (point_groups.py:123-151)."""
    mats = []
    for perm in itertools.permutations(range(3)):
        for sg in itertools.product((1.0, -1.0), repeat=3):
            m = np.zeros((3, 3))
            m[np.arange(3), perm] = sg
            if np.linalg.det(m) > 0.5:
                mats.append(m)
    return np.array(mats)


def _dihedral_group(n):
    """D_n with the n-fold axis along z and a 2-fold axis along x (point_groups.py:36-101: D6 = 622, D4, D3 in
    the a*-along-x frame; for n = 3 the 2-folds lie along the a axes at -30 / 90 / 210 deg, here generated as the
    2-folds at 90 + k 60 deg)."""
    mats = []
    for k in range(n):
        t = 2 * np.pi * k / n
        mats.append(np.array([[np.cos(t), -np.sin(t), 0.0], [np.sin(t), np.cos(t), 0.0], [0.0, 0.0, 1.0]]))
    phi0 = 0.0 if n != 3 else np.pi / 2
    for k in range(n):
        f = phi0 + np.pi * k / n
        mats.append(np.array([[np.cos(2 * f), np.sin(2 * f), 0.0], [np.sin(2 * f), -np.cos(2 * f), 0.0],
                              [0.0, 0.0, -1.0]]))
    return np.array(mats)


def _cyclic_group(n):
    mats = []
    for k in range(n):
        t = 2 * np.pi * k / n
        mats.append(np.array([[np.cos(t), -np.sin(t), 0.0], [np.sin(t), np.cos(t), 0.0], [0.0, 0.0, 1.0]]))
    return np.array(mats)


POINT_GROUPS = {
    "triclinic": lambda: np.eye(3)[None],
    "monoclinic": lambda: _cyclic_group(2),
    "orthorhombic": lambda: _dihedral_group(2),
    "tetragonal": lambda: _dihedral_group(4),
    "trigonal": lambda: _dihedral_group(3),
    "hexagonal": lambda: _dihedral_group(6),
    "cubic": _cube_group,
}


def point_group(system: str) -> np.ndarray:
    """(G, 3, 3) proper rotations of the crystal system's Laue group in the crystal Cartesian frame."""
    s = system.lower()
    if "rhombohedral" in s:
        s = "trigonal"
    if s not in POINT_GROUPS:
        raise ValueError(f"Unknown crystal system '{system}'. Must be one of {list(POINT_GROUPS)}")
    return POINT_GROUPS[s]()


def lattice_matrix_from_params(a, b, c, alpha, beta, gamma):
    """Direct lattice matrix (columns = lattice vectors; material.py:72-105): a along x for 90-degree cells, else the
    Busing-Levy frame (a* along x, c along z), the frame of ImageD11's orientation matrices."""
    if alpha == 90.0 and beta == 90.0 and gamma == 90.0:
        return np.diag([float(a), float(b), float(c)])
    ca, cb, cg = (np.cos(np.deg2rad(t)) for t in (alpha, beta, gamma))
    G = np.array([[a * a, a * b * cg, a * c * cb], [a * b * cg, b * b, b * c * ca], [a * c * cb, b * c * ca, c * c]])
    Gi = np.linalg.inv(G)
    ast, bst, cst = np.sqrt(np.diag(Gi))
    cb_s = Gi[0, 2] / (ast * cst)
    cg_s = Gi[0, 1] / (ast * bst)
    Bm = np.array([[ast, bst * cg_s, cst * cb_s],
                   [0.0, bst * np.sqrt(1 - cg_s ** 2), -cst * np.sqrt(1 - cb_s ** 2) * ca],
                   [0.0, 0.0, 1.0 / c]])
    return np.linalg.inv(Bm).T


def _params_from_rows(M):
    M = np.asarray(M, float)
    a, b, c = np.linalg.norm(M, axis=1)
    ang = lambda u, v: float(np.degrees(np.arccos(np.clip(u @ v / (np.linalg.norm(u) * np.linalg.norm(v)), -1, 1))))
    return a, b, c, ang(M[1], M[2]), ang(M[0], M[2]), ang(M[0], M[1])


def _infer_system(a, b, c, alpha, beta, gamma):
    eq = lambda x, y: abs(x - y) < 1e-4 * max(1.0, abs(y))
    all90 = eq(alpha, 90) and eq(beta, 90) and eq(gamma, 90)
    if eq(a, b) and eq(b, c) and all90:
        return "cubic"
    if eq(a, b) and eq(alpha, 90) and eq(beta, 90) and eq(gamma, 120):
        return "hexagonal"
    if eq(a, b) and all90:
        return "tetragonal"
    if eq(a, b) and eq(b, c) and eq(alpha, beta) and eq(beta, gamma):
        return "trigonal"
    if all90:
        return "orthorhombic"
    if eq(alpha, 90) and eq(gamma, 90):
        return "monoclinic"
    return "triclinic"


class Material:
    """A crystalline phase: lattice, reflection table (hkl family, multiplicity, d, two-theta), h vectors and the
    point-group matrices (the reference's Material as the single-phase operator uses it)."""

    def __init__(self, name: str = ""):
        self.name = name
        self.crystal_system = None
        self.lattice_params = None
        self._A = None
        self._B = None
        self.reflections = None
        self.hkl = self.h_vecs = self.h_vecs_normed = None
        self.point_group_matrices = None
        self.num_sym_ops = None
        self.wavelength_A = None

    @classmethod
    def from_lattice_parameters(cls, *, a=None, b=None, c=None, alpha=90.0, beta=90.0, gamma=90.0,
                                lattice_matrix=None, lattice_matrix_kind="direct", reciprocal_has_2pi=False,
                                symmetry_group=None, crystal_system=None, energy_kev=None, wavelength_A=None,
                                energy_keV=None, min_two_theta=None, max_two_theta=None, q_min=None, q_max=None,
                                hkl_list=None, name="custom_lattice", **_ignored):
        """material.py:639-755 without space-group tables: explicit hkl_list, or every reflection in a q range
        (from q_min/q_max or a two-theta range with the wavelength)."""
        mat = cls(name=name)
        if lattice_matrix is not None:
            M = np.asarray(lattice_matrix, float)
            if M.shape != (3, 3):
                raise ValueError(f"lattice_matrix must be (3, 3), got {M.shape}")
            if lattice_matrix_kind == "reciprocal":
                rec = M / (2 * np.pi) if reciprocal_has_2pi else M
                M = np.linalg.inv(rec).T
            elif lattice_matrix_kind != "direct":
                raise ValueError(f"Unknown lattice_matrix_kind: {lattice_matrix_kind!r}")
            prm = _params_from_rows(M)
        elif a is not None:
            prm = (float(a), float(b) if b is not None else float(a), float(c) if c is not None else float(a),
                   float(alpha), float(beta), float(gamma))
        else:
            raise ValueError("Provide 'lattice_matrix' or 'a' (and optionally b, c, angles).")
        mat.lattice_params = dict(zip(("a", "b", "c", "alpha", "beta", "gamma"), prm))
        mat._A = lattice_matrix_from_params(*prm)
        mat._B = 2.0 * np.pi * np.linalg.inv(mat._A).T
        sys_ = crystal_system or symmetry_group
        mat.crystal_system = sys_.lower() if sys_ else _infer_system(*prm)
        lam = None
        if wavelength_A is not None:
            lam = float(wavelength_A)
        elif energy_keV is not None or energy_kev is not None:
            lam = 12.39842 / float(energy_keV if energy_keV is not None else energy_kev)
        mat.wavelength_A = lam
        if hkl_list is None:
            if q_min is not None and q_max is not None:
                qlo, qhi = float(q_min), float(q_max)
            elif min_two_theta is not None and max_two_theta is not None and lam is not None:
                qlo = 4 * np.pi / lam * np.sin(float(min_two_theta) / 2)
                qhi = 4 * np.pi / lam * np.sin(float(max_two_theta) / 2)
            else:
                raise ValueError("Insufficient input: provide q_min+q_max, or (wavelength/energy + two-theta range), "
                                 "or an explicit hkl_list.")
        else:
            qlo, qhi = 0.0, np.inf
        mat.compute_reflections(q_min=qlo, q_max=qhi, wavelength_A=lam, hkl_list=hkl_list)
        return mat

    def _hkl_ops(self):
        """Point-group rotations acting on Miller indices: B^-1 R B (integers for the lattice's own group)."""
        Bi = np.linalg.inv(self._B)
        return [np.rint(Bi @ r @ self._B).astype(np.int64) for r in point_group(self.crystal_system)]

    def compute_reflections(self, *, q_min=0.0, q_max, wavelength_A=None, hkl_list=None):
        """Reflection families (material.py:1000-1174): equivalents under the point group and Friedel pairs,
        representative = the smallest equivalent, multiplicity = their number; sorted by two-theta."""
        B = self._B
        if hkl_list is not None:
            raw = [tuple(int(v) for v in h) for h in hkl_list]
        else:
            if not np.isfinite(q_max) or q_max <= 0:
                raise ValueError("q_max must be a positive finite number.")
            lens = np.linalg.norm(B, axis=0)
            n = [int(np.ceil(q_max / L)) + 1 for L in lens]
            grid = np.stack(np.meshgrid(*[np.arange(-m, m + 1) for m in n], indexing="ij"), -1).reshape(-1, 3)
            grid = grid[np.any(grid != 0, axis=1)]
            q = np.linalg.norm(B @ grid.T, axis=0)
            raw = [tuple(int(v) for v in h) for h in grid[(q >= q_min) & (q <= q_max)]]
        if not raw:
            raise ValueError("No reflections found in the specified q range.")
        ops = self._hkl_ops()
        fams, mults, seen = [], [], set()
        for h in raw:
            if h in seen:
                continue
            eq = []
            for R in ops:
                w = tuple(int(v) for v in R @ np.asarray(h))
                for cand in (w, tuple(-v for v in w)):
                    if cand not in eq:
                        eq.append(cand)
            seen.update(eq)
            fams.append(min(eq))
            mults.append(len(eq))
        hkl = np.array(fams, dtype=np.int32)
        q = np.linalg.norm(B @ hkl.T, axis=0)
        d = 2 * np.pi / q
        if wavelength_A is not None:
            st = q * float(wavelength_A) / (4 * np.pi)
            ok = st <= 1.0
            tth = np.where(ok, 2 * np.arcsin(np.clip(st, 0, 1)), 0.0)
        else:
            ok = np.ones(len(hkl), bool)
            tth = np.zeros(len(hkl))
        hkl, mults, d, tth = hkl[ok], np.asarray(mults)[ok], d[ok], tth[ok]
        order = np.argsort(tth if wavelength_A is not None else -d, kind="stable")
        r = np.empty(len(order), dtype=_REFLECTION_DTYPE)
        r["hkl"], r["multiplicity"], r["d_spacing"] = hkl[order], mults[order], d[order]
        r["two_theta"], r["intensity"], r["sf_squared"] = tth[order], np.nan, np.nan
        self.reflections = r
        self.hkl = r["hkl"].copy()
        h = (B @ r["hkl"].astype(float).T).T
        self.h_vecs = h
        self.h_vecs_normed = h / (np.linalg.norm(h, axis=1, keepdims=True) + 1e-12)
        pg = point_group(self.crystal_system)
        self.point_group_matrices = pg.reshape(len(pg), 9).astype(np.float32)
        self.num_sym_ops = len(pg)

    def multiplicities(self):
        return self.reflections["multiplicity"]

    def two_theta(self):
        return self.reflections["two_theta"]

    def d_spacings(self):
        return self.reflections["d_spacing"]

    def intensities(self):
        return self.reflections["intensity"]

    def __len__(self):
        return 0 if self.reflections is None else len(self.reflections)


def group_reflections_into_rings(material, rtol=1e-6):
    """Reflections of equal two-theta form one data channel, sorted by two-theta (intensity.py:156-169)."""
    tt = np.asarray(material.reflections["two_theta"], np.float64)
    rings = []
    for i in np.argsort(tt, kind="stable"):
        if rings and np.isclose(tt[i], tt[rings[-1][0]], rtol=rtol, atol=0.0):
            rings[-1].append(int(i))
        else:
            rings.append([int(i)])
    return rings
