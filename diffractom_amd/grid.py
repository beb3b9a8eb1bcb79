"""Orientation grids: K rotation matrices (crystal -> sample) and their Gaussian kernel widths (the reference's
utils/grid.py Grid and MatrixGrid, single level)."""
from __future__ import annotations

import numpy as np
from scipy.spatial import cKDTree
from scipy.spatial.transform import Rotation

from .crystallography import point_group


class MatrixGrid:
    """K rotation matrices and their kernel widths as arrays (utils/grid.py:610-645)."""

    def __init__(self, R_mats, sigma):
        R_mats = np.asarray(R_mats, dtype=np.float64)
        if R_mats.ndim != 3 or R_mats.shape[1:] != (3, 3):
            raise ValueError("R_mats must have shape (N, 3, 3)")
        self.matrices = R_mats
        self.sigmas = np.broadcast_to(np.asarray(sigma, dtype=np.float64), (len(R_mats),)).copy()
        self.sigma_levels = [float(self.sigmas[0])] if len(R_mats) else []

    def __len__(self):
        return len(self.matrices)

    def leaf_arrays(self):
        return self.matrices, self.sigmas

    def active_leaf_nodes(self):
        return range(len(self.matrices))

    def nodes_at_level(self, level):
        if level != 0:
            raise ValueError("single-level grid")
        return range(len(self.matrices))

    def rotations_at_level(self, level, active_only=True):
        if level != 0:
            raise ValueError("single-level grid")
        return [Rotation.from_matrix(self.matrices)]


class Grid(MatrixGrid):
    """A single-level grid built from rotations (utils/grid.py:238-378, 499-606). The matrices are those of scipy
    Rotation objects (orthonormalised through their quaternions), as the reference's per-node Rotations give them."""

    @classmethod
    def from_rotation_matrices(cls, R_mats, sigma):
        R_mats = np.asarray(R_mats)
        if R_mats.ndim != 3 or R_mats.shape[1:] != (3, 3):
            raise ValueError("R_mats must have shape (N, 3, 3)")
        return cls(Rotation.from_matrix(R_mats).as_matrix(), sigma)

    @classmethod
    def from_random_fundamental_zone(cls, n_orientations, symmetry, sigma, random_state=None):
        """Uniform random rotations (scipy Rotation.random, global numpy state unless random_state is given), each
        replaced by the symmetry equivalent R g with the largest trace (grid.py:295-378)."""
        pg = point_group(symmetry)
        Rr = Rotation.random(n_orientations, random_state=random_state).as_matrix()
        prod = np.einsum("nij,gjk->ngik", Rr, pg)
        best = np.argmax(np.einsum("ngii->ng", prod), axis=1)
        Rfz = prod[np.arange(n_orientations), best]
        return cls(Rotation.from_matrix(Rfz).as_matrix(), sigma)

    def prune_close_orientations(self, theta_deg, target=None, symmetry=None):
        """Greedy pruning of orientations closer than theta_deg to a kept one, with a KD-tree on unit quaternions
        (grid.py:499-606; symmetry None: plain rotation distances, antipodal copies mark rows never visited)."""
        N = len(self.matrices)
        if N == 0:
            return 0
        R_all = Rotation.from_matrix(self.matrices)
        q = R_all.as_quat()
        q /= np.linalg.norm(q, axis=1, keepdims=True)
        if symmetry is not None:
            q_full = np.concatenate([(R_all * Rotation.from_matrix(g)).as_quat() for g in point_group(symmetry)])
        else:
            q_full = q
        q_full = np.vstack([q_full, -q_full])
        radius = 2.0 * np.sin(np.deg2rad(theta_deg) / 4.0)
        kd = cKDTree(q_full)
        used = np.zeros(N if symmetry is not None else 2 * N, dtype=bool)
        keep = np.zeros(N, dtype=bool)
        kept = 0
        for i in range(N):
            if used[i]:
                continue
            keep[i] = True
            kept += 1
            nb = np.asarray(kd.query_ball_point(q[i], r=radius), dtype=np.int64)
            used[nb % N if symmetry is not None else nb] = True
            if target is not None and kept >= target:
                break
        self.matrices = self.matrices[keep]
        self.sigmas = self.sigmas[keep]
        return len(self.matrices)
