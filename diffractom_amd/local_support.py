"""Local bases: where each orientation of the basis may carry a coefficient (operators/local_support.py:1-283).

A LocalSupport is a set of allowed (orientation k, voxel (iy, ix)) pairs for K orientations on an Ny x Nx grid, kept
sorted by (k, iy, ix) and unique. build_patches groups the pairs of every orientation into rectangular boxes
("patches"); the local-basis operator stores a patch's coefficients as a small dense row-major image, zero where no
pair is.

Things that fix the result and therefore follow the reference: the pair order, the disc of dilate (offsets with
dy^2 + dx^2 <= r^2 + 1/2), the rule that splits an orientation into clusters (voxels that fill less than
min_box_fill of its bounding box; clusters = 8-connected components after growing every voxel by a
(2 merge_gap + 1)^2 square), that overlapping cluster boxes are merged until none overlap, and the patch order
(orientation, then the boxes sorted by (top, bottom, left, right)), which defines the flat coefficient layout.
"""
from __future__ import annotations

import numpy as np
from scipy import ndimage
from scipy.spatial import cKDTree
from scipy.spatial.transform import Rotation

from .crystallography import point_group


def disc_offsets(radius):
    """(dy, dx) of the voxels within `radius` of a voxel: dy^2 + dx^2 <= radius^2 + 1/2 (radius 1: a cross)."""
    r = int(np.ceil(radius))
    dy, dx = np.meshgrid(np.arange(-r, r + 1), np.arange(-r, r + 1), indexing="ij")
    inside = dy * dy + dx * dx <= radius * radius + 0.5
    return dy[inside], dx[inside]


def _symmetry_matrices(symmetry):
    """(G, 3, 3) proper rotations of a crystal system name, a Material or a scipy Rotation of the point group."""
    if isinstance(symmetry, str):
        return point_group(symmetry)
    if isinstance(symmetry, Rotation):
        return symmetry.as_matrix().reshape(-1, 3, 3)
    name = getattr(symmetry, "crystal_system", None)
    if not isinstance(name, str):
        raise TypeError("symmetry must be a crystal system name, a Material or a scipy Rotation")
    return point_group(name)


class LocalSupport:
    """Allowed (k, iy, ix) pairs of a basis of K orientations on an (Ny, Nx) grid, sorted by (k, iy, ix), unique."""

    def __init__(self, k, iy, ix, K, shape):
        self.K = int(K)
        self.Ny, self.Nx = (int(s) for s in shape)
        k, iy, ix = (np.asarray(a, dtype=np.int64).reshape(-1) for a in (k, iy, ix))
        if not (k.size == iy.size == ix.size):
            raise ValueError("k, iy and ix must have the same length")
        bad = (k < 0) | (k >= self.K) | (iy < 0) | (iy >= self.Ny) | (ix < 0) | (ix >= self.Nx)
        if bad.any():
            raise ValueError(f"{int(bad.sum())} pairs lie outside the basis (K = {self.K}) or the grid "
                             f"({self.Ny} x {self.Nx})")
        key = np.unique((k * self.Ny + iy) * self.Nx + ix)
        npix = self.Ny * self.Nx
        self.k = key // npix
        self.iy = (key % npix) // self.Nx
        self.ix = key % self.Nx

    # ---------------------------------------------------------------- construction
    @classmethod
    def from_mask(cls, mask):
        """From a (K, Ny, Nx) boolean array."""
        m = np.asarray(mask, dtype=bool)
        if m.ndim != 3:
            raise ValueError("mask must have shape (K, Ny, Nx)")
        k, iy, ix = np.nonzero(m)
        return cls(k, iy, ix, m.shape[0], m.shape[1:])

    @classmethod
    def everywhere(cls, K, shape):
        """Every orientation in every voxel (the global basis)."""
        return cls.from_mask(np.ones((int(K),) + tuple(int(s) for s in shape), dtype=bool))

    @classmethod
    def from_candidates(cls, basis, iy, ix, orientations, symmetry, shape, tol_deg=1.0, dilate=2):
        """Orientation k is allowed in voxel (iy, ix) when a candidate orientation found there lies within tol_deg of
        it (crystal symmetry included); then every orientation's voxels grow by a disc of `dilate` voxels.

        basis (K, 3, 3) and orientations (N, 3, 3): rotation matrices (crystal -> sample); iy, ix (N,): the voxel of
        every candidate; symmetry: crystal system name, Material or scipy Rotation of the point group."""
        basis = np.asarray(basis, dtype=np.float64).reshape(-1, 3, 3)
        K = len(basis)
        sym = _symmetry_matrices(symmetry)
        # the symmetry equivalents U g of every basis orientation as unit quaternions, with both signs: the chord
        # |q - q'| of the nearer sign is 2 sin(angle / 4)
        q_eq = np.concatenate([Rotation.from_matrix(basis @ g).as_quat() for g in sym])
        q_eq = np.vstack([q_eq, -q_eq])
        tree = cKDTree(q_eq)
        qc = Rotation.from_matrix(np.asarray(orientations, dtype=np.float64).reshape(-1, 3, 3)).as_quat()
        hits = tree.query_ball_point(qc, 2.0 * np.sin(np.deg2rad(tol_deg) / 4.0))
        counts = np.array([len(h) for h in hits], dtype=np.int64)
        ks = (np.concatenate([np.asarray(h, np.int64) for h in hits]) % K) if counts.sum() else \
            np.zeros(0, np.int64)
        sup = cls(ks, np.repeat(np.asarray(iy, np.int64), counts), np.repeat(np.asarray(ix, np.int64), counts), K,
                  shape)
        return sup.dilate(dilate) if dilate > 0 else sup

    @classmethod
    def from_coefficients(cls, coeffs, top=3, rel=0.0, dilate=2, mask=None, min_voxel=0.0, chunk_elems=50_000_000):
        """From a (K, Ny, Nx) fit: orientation k is allowed in a voxel when its coefficient there is positive, among
        the `top` largest of the voxel and at least rel x the voxel's largest, in voxels whose largest coefficient is
        at least min_voxel x the largest overall; then the voxels grow by a disc of `dilate` voxels and are cut to the
        (Ny, Nx) mask if one is given."""
        c = np.asarray(coeffs)
        if c.ndim != 3:
            raise ValueError("coeffs must have shape (K, Ny, Nx)")
        K, Ny, Nx = c.shape
        top = max(1, min(int(top), K))
        flat = c.reshape(K, Ny * Nx)
        floor = float(min_voxel) * float(flat.max()) if min_voxel else 0.0
        step = max(1, int(chunk_elems) // K)
        ks, ps = [], []
        for p0 in range(0, Ny * Nx, step):
            blk = np.asarray(flat[:, p0:p0 + step], dtype=np.float32)
            if top < K:
                idx = np.argpartition(-blk, top - 1, axis=0)[:top]
            else:
                idx = np.broadcast_to(np.arange(K)[:, None], blk.shape)
            val = np.take_along_axis(blk, idx, axis=0)
            vmax = blk.max(axis=0)
            keep = (val > 0) & (val >= rel * vmax) & (vmax >= floor)
            ks.append(idx[keep].astype(np.int64))
            ps.append(np.nonzero(keep)[1].astype(np.int64) + p0)
        k = np.concatenate(ks) if ks else np.zeros(0, np.int64)
        p = np.concatenate(ps) if ps else np.zeros(0, np.int64)
        sup = cls(k, p // Nx, p % Nx, K, (Ny, Nx))
        if dilate > 0:
            sup = sup.dilate(dilate)
        return sup.restrict(mask) if mask is not None else sup

    # ---------------------------------------------------------------- set operations
    def dilate(self, radius):
        """Every orientation's voxels grown by a disc of `radius` voxels (cut at the grid's edges)."""
        dy, dx = disc_offsets(radius)
        iy = (self.iy[:, None] + dy[None]).reshape(-1)
        ix = (self.ix[:, None] + dx[None]).reshape(-1)
        k = np.repeat(self.k, dy.size)
        ok = (iy >= 0) & (iy < self.Ny) & (ix >= 0) & (ix < self.Nx)
        return LocalSupport(k[ok], iy[ok], ix[ok], self.K, (self.Ny, self.Nx))

    def restrict(self, mask):
        """Only the voxels where the (Ny, Nx) mask is true."""
        m = np.asarray(mask, dtype=bool)
        if m.shape != (self.Ny, self.Nx):
            raise ValueError(f"mask shape {m.shape} != {(self.Ny, self.Nx)}")
        keep = m[self.iy, self.ix]
        return LocalSupport(self.k[keep], self.iy[keep], self.ix[keep], self.K, (self.Ny, self.Nx))

    def union(self, other):
        if (other.K, other.Ny, other.Nx) != (self.K, self.Ny, self.Nx):
            raise ValueError("supports of different bases or grids")
        return LocalSupport(np.concatenate([self.k, other.k]), np.concatenate([self.iy, other.iy]),
                            np.concatenate([self.ix, other.ix]), self.K, (self.Ny, self.Nx))

    def to_mask(self):
        m = np.zeros((self.K, self.Ny, self.Nx), dtype=bool)
        m[self.k, self.iy, self.ix] = True
        return m

    @property
    def n_pairs(self):
        return int(self.k.size)

    @property
    def fill(self):
        """Share of the K x Ny x Nx coefficients that are allowed."""
        return self.n_pairs / float(self.K * self.Ny * self.Nx)

    def orientations_used(self):
        return np.unique(self.k)

    def per_voxel(self):
        """(Ny, Nx): allowed orientations per voxel."""
        return np.bincount(self.iy * self.Nx + self.ix, minlength=self.Ny * self.Nx).reshape(self.Ny, self.Nx)

    def __repr__(self):
        return (f"LocalSupport({self.n_pairs} pairs, {self.orientations_used().size} of {self.K} orientations, "
                f"{self.Ny} x {self.Nx} voxels, fill {100 * self.fill:.2f} %)")


# ---------------------------------------------------------------------- patches
def merge_overlapping(boxes):
    """Boxes (y0, y1, x0, x1), inclusive, merged into their bounding boxes until no two overlap: the connected
    components of the overlap relation, repeated while a merged box meets another one."""
    boxes = np.asarray(boxes, dtype=np.int64).reshape(-1, 4)
    while len(boxes) > 1:
        a, b = boxes[:, None, :], boxes[None, :, :]
        meet = (a[..., 0] <= b[..., 1]) & (b[..., 0] <= a[..., 1]) & (a[..., 2] <= b[..., 3]) & (b[..., 2] <= a[..., 3])
        n_comp, comp = _components(meet)
        if n_comp == len(boxes):
            break
        out = np.empty((n_comp, 4), dtype=np.int64)
        for c in range(n_comp):
            sel = boxes[comp == c]
            out[c] = (sel[:, 0].min(), sel[:, 1].max(), sel[:, 2].min(), sel[:, 3].max())
        boxes = out
    return [tuple(int(v) for v in bx) for bx in boxes]


def _components(adj):
    """Connected components of a symmetric boolean adjacency matrix (labels 0 .. n - 1)."""
    n = len(adj)
    lab = np.full(n, -1, dtype=np.int64)
    c = 0
    for s in range(n):
        if lab[s] >= 0:
            continue
        stack = [s]
        lab[s] = c
        while stack:
            i = stack.pop()
            for j in np.flatnonzero(adj[i] & (lab < 0)):
                lab[j] = c
                stack.append(int(j))
        c += 1
    return c, lab


def _orientation_boxes(iy, ix, split, merge_gap, min_box_fill):
    """The boxes (y0, y1, x0, x1) of one orientation's voxels, sorted."""
    y0, y1, x0, x1 = int(iy.min()), int(iy.max()), int(ix.min()), int(ix.max())
    area = (y1 - y0 + 1) * (x1 - x0 + 1)
    if not split or iy.size >= min_box_fill * area:
        return [(y0, y1, x0, x1)]
    occ = np.zeros((y1 - y0 + 1, x1 - x0 + 1), dtype=bool)
    occ[iy - y0, ix - x0] = True
    if merge_gap > 0:
        occ = ndimage.binary_dilation(occ, structure=np.ones((2 * merge_gap + 1,) * 2, dtype=bool))
    lab, n = ndimage.label(occ, structure=np.ones((3, 3), dtype=bool))
    lv = lab[iy - y0, ix - x0]
    boxes = []
    for c in np.unique(lv):
        s = lv == c
        boxes.append((int(iy[s].min()), int(iy[s].max()), int(ix[s].min()), int(ix[s].max())))
    return sorted(merge_overlapping(boxes))


def build_patches(k, iy, ix, K, split=True, merge_gap=2, min_box_fill=0.5):
    """Rectangular patches of sorted, unique (k, iy, ix) pairs (operators/local_support.py:215-283): one box per
    orientation, or, with split, one per cluster of an orientation whose voxels fill less than min_box_fill of its
    bounding box. Returns dict(pk, y0, x0, h, w, coff, n_coef, mask, pair_index): the orientation, top-left corner
    and size of every patch, the offset of its row-major coefficients in the flat vector, the total length, a uint8
    flag per coefficient (1 where a pair is) and the flat index of every pair (in the pairs' order)."""
    k, iy, ix = (np.asarray(a, dtype=np.int64).reshape(-1) for a in (k, iy, ix))
    n = k.size
    if n == 0:
        raise ValueError("empty support")
    bounds = np.flatnonzero(np.diff(k)) + 1
    starts = np.concatenate([[0], bounds])
    ends = np.concatenate([bounds, [n]])
    pk, py0, px0, ph, pw = [], [], [], [], []
    patch_of = np.empty(n, dtype=np.int64)
    for s, e in zip(starts, ends):
        yy, xx = iy[s:e], ix[s:e]
        for (b0, b1, c0, c1) in _orientation_boxes(yy, xx, split, int(merge_gap), float(min_box_fill)):
            inside = (yy >= b0) & (yy <= b1) & (xx >= c0) & (xx <= c1)
            patch_of[s:e][inside] = len(pk)
            pk.append(int(k[s]))
            py0.append(b0)
            px0.append(c0)
            ph.append(b1 - b0 + 1)
            pw.append(c1 - c0 + 1)
    pk, py0, px0, ph, pw = (np.asarray(a, dtype=np.int64) for a in (pk, py0, px0, ph, pw))
    size = ph * pw
    coff = np.concatenate([[0], np.cumsum(size)[:-1]]).astype(np.int64)
    n_coef = int(size.sum())
    pair_index = coff[patch_of] + (iy - py0[patch_of]) * pw[patch_of] + (ix - px0[patch_of])
    mask = np.zeros(n_coef, dtype=np.uint8)
    mask[pair_index] = 1
    if int(mask.sum()) != n:
        raise RuntimeError("patches of one orientation overlap")
    return dict(pk=pk, y0=py0, x0=px0, h=ph, w=pw, coff=coff, n_coef=n_coef, mask=mask, pair_index=pair_index)


def morton_code(y, x):
    """Z-order code of non-negative integer coordinates below 2^16: the bits of y and x interleaved (y higher)."""
    def spread(v):
        v = np.asarray(v, dtype=np.uint64) & np.uint64(0xFFFF)
        out = np.zeros_like(v)
        for b in range(16):
            out |= ((v >> np.uint64(b)) & np.uint64(1)) << np.uint64(2 * b)
        return out
    return (spread(y) << np.uint64(1)) | spread(x)


def patch_neighbours(patches):
    """(right, down, left, up) int32 index arrays of the flat coefficients of build_patches' layout: the neighbour
    within the coefficient's patch, -1 at the box's edges (spdhg_poisson.py:982-1000)."""
    h, w = patches["h"].astype(np.int64), patches["w"].astype(np.int64)
    size = h * w
    n = int(size.sum())
    j = np.arange(n, dtype=np.int64)
    loc = j - np.repeat(patches["coff"].astype(np.int64), size)
    ww, hh = np.repeat(w, size), np.repeat(h, size)
    row, col = loc // ww, loc % ww
    right = np.where(col < ww - 1, j + 1, -1)
    down = np.where(row < hh - 1, j + ww, -1)
    left = np.where(col > 0, j - 1, -1)
    up = np.where(row > 0, j - ww, -1)
    return [a.astype(np.int32) for a in (right, down, left, up)]
