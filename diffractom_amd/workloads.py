"""Synthetic workloads of the benchmark configurations: materials, scan geometry, orientation grids, Voronoi
ground truths and Poisson data. Every generator is seeded and deterministic.

  al_c0  Al (Fm-3m, a 4.05 A, lambda 0.2 A), 5 rings, 180 omega over 180 deg, My 103, 180 eta, 99 x 99 pixels,
         K = 2000 random fundamental-zone orientations, sigma 2 deg; 40 Voronoi grains in a disc of radius 45 px
         (one orientation each, coefficient 1); Poisson counts at a mean of 1000; SPDHG M = 10, 200 steps.
  al_c1  The same scan with a uniform cubic fundamental-zone grid at 3 deg spacing (random orientations thinned to
         0.6 x 3 deg: K = 29713), sigma 2 deg;
         80 grains with coefficients U(0.5, 1.5); SPDHG M = 10, 300 steps.
  al_local  The al_c1 scan, grid and truth on a local basis: every voxel allows the orientations within 3.5 deg of
         its grain's (grown by 2 voxels).
"""
from __future__ import annotations

import numpy as np
from scipy.spatial import cKDTree
from scipy.spatial.transform import Rotation

from .crystallography import Material, point_group
from .grid import MatrixGrid

AL_RINGS = [[1, 1, 1], [2, 0, 0], [2, 2, 0], [3, 1, 1], [2, 2, 2]]
TI_RINGS = [[1, 0, 0], [0, 0, 2], [1, 0, 1], [1, 0, 2], [1, 1, 0], [1, 0, 3], [2, 0, 0], [1, 1, 2]]


def al_material(wavelength=0.2):
    return Material.from_lattice_parameters(a=4.05, b=4.05, c=4.05, alpha=90, beta=90, gamma=90,
                                            symmetry_group="cubic", wavelength_A=wavelength, hkl_list=AL_RINGS)


def ti_material(wavelength=0.2):
    return Material.from_lattice_parameters(a=2.95, b=2.95, c=4.68, alpha=90, beta=90, gamma=120,
                                            symmetry_group="hexagonal", wavelength_A=wavelength, hkl_list=TI_RINGS)


def scan_cfg(n, my, n_omega, n_eta, wavelength=0.2, angle_range=(0, 180), cor_offset=0.0):
    """The reference's cfg (README conventions: rotation axis z, beam x, eta 0 along -y, eta 90 along -z)."""
    return {"Nx": n, "Ny": n, "My": my, "N_Omega": n_omega, "angle_range": list(angle_range),
            "cor_offset": cor_offset, "N_eta": n_eta, "eta_angle_range": [0, 360], "energy": 12.398 / wavelength,
            "k_direction_0": [0, 0, 1], "p_direction_0": [1, 0, 0], "detector_direction_origin": [0, -1, 0],
            "detector_direction_positive_90": [0, 0, -1]}


def to_fundamental_zone(R, symmetry):
    """Each rotation replaced by the equivalent R g with the largest trace (utils/grid.py:341-358)."""
    pg = point_group(symmetry)
    prod = np.einsum("nij,gjk->ngik", R, pg)
    best = np.argmax(np.einsum("ngii->ng", prod), axis=1)
    return prod[np.arange(len(R)), best]


def random_fz(K, symmetry, seed):
    """K uniform random orientations (normalised Gaussian quaternions from default_rng(seed)) in the FZ."""
    q = np.random.default_rng(seed).standard_normal((K, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    return to_fundamental_zone(Rotation.from_quat(q).as_matrix(), symmetry)


def uniform_fz(spacing_deg, symmetry, seed=0, thin=0.7):
    """Near-uniform FZ orientations: 3 x (SO(3) volume / spacing^3 / |G|) random orientations, thinned greedily to a
    minimum misorientation of thin x spacing (the "random" method of odf_basis.py:117-139 thins at 0.7)."""
    pg = point_group(symmetry)
    n = int(3 * 8 * np.pi ** 2 / np.deg2rad(spacing_deg) ** 3 / len(pg))
    R = random_fz(n, symmetry, seed)
    rot = Rotation.from_matrix(R)
    q = rot.as_quat()
    eq = np.concatenate([(rot * Rotation.from_matrix(g)).as_quat() for g in pg])
    eq = np.vstack([eq, -eq])
    tree = cKDTree(eq)
    radius = 2.0 * np.sin(np.deg2rad(thin * spacing_deg) / 4.0)
    used = np.zeros(n, bool)
    keep = np.zeros(n, bool)
    for i in range(n):
        if used[i]:
            continue
        keep[i] = True
        used[np.asarray(tree.query_ball_point(q[i], r=radius), dtype=np.int64) % n] = True
    return R[keep]


def voronoi_truth(K, n, n_grains, radius, seed, coeff=None, orientations=None):
    """(K, n, n) coefficients: n_grains Voronoi cells of seeds uniform in a centred disc of `radius` pixels; every
    pixel of the disc gets its cell's orientation with coefficient 1 (coeff None) or U(coeff) per grain."""
    k, iy, ix, v = voronoi_pairs(K, n, n_grains, radius, seed, coeff, orientations)
    x = np.zeros((K, n, n), np.float32)
    x[k, iy, ix] = v
    return x


def voronoi_pairs(K, n, n_grains, radius, seed, coeff=None, orientations=None):
    """The non-zeros of voronoi_truth as (orientation, row, column, value) arrays, without the (K, n, n) array."""
    rng = np.random.default_rng(seed)
    c = (n - 1) / 2.0
    r = radius * np.sqrt(rng.random(n_grains))
    t = rng.uniform(0, 2 * np.pi, n_grains)
    sx, sy = c + r * np.cos(t), c + r * np.sin(t)
    ori = rng.choice(K, n_grains, replace=False) if orientations is None else np.asarray(orientations)
    val = np.ones(n_grains) if coeff is None else rng.uniform(coeff[0], coeff[1], n_grains)
    yy, xx = np.mgrid[:n, :n]
    inside = (xx - c) ** 2 + (yy - c) ** 2 <= radius ** 2
    d2 = (xx[..., None] - sx) ** 2 + (yy[..., None] - sy) ** 2
    cell = np.argmin(d2, axis=-1)
    iy, ix = np.nonzero(inside)
    return ori[cell[iy, ix]], iy, ix, val[cell[iy, ix]].astype(np.float32)


LOCAL = dict(tol_deg=3.5, dilate=2)


def al_local_problem(precision="bf16_residual", pf_mode="sparse", max_gb=1024.0):
    """The local-basis workload: the al_c1 scan and 3-degree grid (K = 29713) and the al_c1 Voronoi truth (80 grains
    in a disc of radius 45 px); every voxel allows the basis orientations within 3.5 deg of its grain's orientation
    (what point-by-point indexing finds), grown by 2 voxels (LocalSupport.from_candidates). Counts: the local
    forward projection of the truth at a mean of 1000. Returns (local operator, truth (n_coef,) float32, counts,
    parameters)."""
    from .local_operator import LocalSinglePhaseForwardOperator
    from .local_support import LocalSupport
    p = CONFIGS["al_c1"]
    grid, mat, cfg = make_grid("al_c1")
    K, n = len(grid), p["n"]
    k, iy, ix, v = voronoi_pairs(K, n, p["grains"], p["radius"], p["seed"], p["coeff"])
    sup = LocalSupport.from_candidates(grid.matrices, iy, ix, grid.matrices[k], "cubic", (n, n),
                                       tol_deg=LOCAL["tol_deg"], dilate=LOCAL["dilate"])
    op = LocalSinglePhaseForwardOperator(cfg, mat, grid, max_gb, support=sup, normalized=True, pf_mode=pf_mode,
                                         pf_precision=precision)
    s = op.support
    keys = (s.k * n + s.iy) * n + s.ix
    want = (k * n + iy) * n + ix
    at = np.searchsorted(keys, want)
    if not np.array_equal(keys[np.minimum(at, keys.size - 1)], want):
        raise RuntimeError("the truth lies outside the local support")
    xt = np.zeros(op.n_coef, np.float32)
    xt[op._patches["pair_index"][at]] = v
    b, _ = poisson_counts(op.direct(xt), p["mean_counts"], p["seed"])
    return op, xt, b, {**p, **LOCAL}


def poisson_counts(ax, mean_counts, seed, weights=None):
    """Poisson counts with the prediction scaled to a mean of mean_counts over the counted points."""
    ax = np.asarray(ax, np.float64)
    w = np.ones(ax.size) if weights is None else np.asarray(weights, np.float64).ravel()[np.arange(ax.size) %
                                                                                      np.asarray(weights).size]
    counted = w.reshape(ax.shape) != 0
    s = mean_counts / ax[counted].mean()
    return np.random.default_rng(seed).poisson(s * np.maximum(ax, 0)).astype(np.float32), s


CONFIGS = {
    "al_c0": dict(n=99, my=103, n_omega=180, n_eta=180, K=2000, grid="random", sigma_deg=2.0, grains=40,
                  radius=45, coeff=None, mean_counts=1000, subsets=10, steps=200, seed=0),
    "al_c1": dict(n=99, my=103, n_omega=180, n_eta=180, K=None, grid="uniform3", sigma_deg=2.0, grains=80,
                  radius=45, coeff=(0.5, 1.5), mean_counts=1000, subsets=10, steps=300, seed=1),
}


def make_grid(name, K=None):
    """(MatrixGrid, material, cfg) of a named configuration (K overrides the grid size of a random grid)."""
    p = CONFIGS[name]
    if p["grid"] == "random":
        R = random_fz(K or p["K"], "cubic", p["seed"])
    else:
        R = uniform_fz(3.0, "cubic", seed=p["seed"], thin=0.6)      # 29713 orientations
        if K is not None:
            R = R[:K]
    return MatrixGrid(R, np.deg2rad(p["sigma_deg"])), al_material(), scan_cfg(p["n"], p["my"], p["n_omega"],
                                                                                p["n_eta"])


C3 = dict(n=128, my=131, n_omega=180, n_eta=180, sigma_deg=2.0, slices=64, K=(500, 8000), radius=(15.0, 60.0),
          jitter_deg=1.0, grains_per=50, mean_counts=1000.0, seed=3, subsets=10)


def jittered_grid(K, n_base, jitter_deg, symmetry, seed):
    """K orientations: n_base random fundamental-zone orientations (the grains) and their copies rotated by random
    axes through normally distributed angles of jitter_deg (the first n_base rows are the grains themselves)."""
    from scipy.spatial.transform import Rotation
    rng = np.random.default_rng(seed)
    base = random_fz(n_base, symmetry, int(rng.integers(1 << 31)))
    reps = rng.integers(0, n_base, K - n_base)
    axes = rng.standard_normal((K - n_base, 3))
    axes /= np.linalg.norm(axes, axis=1, keepdims=True)
    ang = np.deg2rad(jitter_deg) * rng.standard_normal(K - n_base)
    J = Rotation.from_rotvec(axes * ang[:, None]).as_matrix()
    return np.concatenate([base, np.einsum("kij,kjl->kil", J, base[reps])])


def c3_slices(n_slices=None, seed=None):
    """The multi-slice workload: slices of 128 x 128 pixels (Al, 180 omega x 131 translations x 180 eta x 5 rings),
    slice i with K_i ~ U(500, 8000) orientations (grains with 1-degree jittered copies), a disc support of radius
    r_i ~ U(15, 60) pixels and a Voronoi truth of its grains inside the disc. Returns (cfg, material, [(grid, support
    mask, truth), ...])."""
    p = C3
    n_slices = p["slices"] if n_slices is None else int(n_slices)
    rng = np.random.default_rng(p["seed"] if seed is None else seed)
    cfg = scan_cfg(p["n"], p["my"], p["n_omega"], p["n_eta"])
    n = p["n"]
    yy, xx = np.mgrid[:n, :n]
    c = (n - 1) / 2.0
    out = []
    for i in range(n_slices):
        K = int(rng.integers(p["K"][0], p["K"][1] + 1))
        r = float(rng.uniform(*p["radius"]))
        n_base = max(1, K // p["grains_per"])
        R = jittered_grid(K, n_base, p["jitter_deg"], "cubic", int(rng.integers(1 << 31)))
        mask = ((xx - c) ** 2 + (yy - c) ** 2 <= r * r).astype(np.uint8)
        n_grains = max(1, min(n_base, int(np.pi * r * r / 150)))
        truth = voronoi_truth(K, n, n_grains, r, int(rng.integers(1 << 31)), coeff=(0.5, 1.5),
                              orientations=rng.choice(n_base, n_grains, replace=False))
        out.append((MatrixGrid(R, np.deg2rad(p["sigma_deg"])), mask, truth))
    return cfg, al_material(), out


C2 = dict(n=256, my=261, n_omega=360, n_eta=360, sigma_deg=2.5, grains=300, fibre=0.1, fibre_deg=10.0,
          mean_counts=500.0, seed=2, subsets=10, radius=120.0)


def c2_problem(device="cuda", max_gb=20.0, pf_mode="sparse", pf_precision="fp32", counts_from=None):
    """The large hexagonal workload: Ti (8 rings), 360 omega over 180 deg x 261 translations x 360 eta, 256 x 256
    pixels, a uniform 3-degree hexagonal fundamental-zone grid (59,440 orientations), sigma 2.5 deg; the PF stored
    as FP32 CSR (pf_mode "sparse", about 48 GB for both forms) or generated per call in fp16 blocks (pf_mode
    "generated", pf_precision "fp16"; the dense BF16 fragments would need about 250 GB). Truth, in the
    orientation-fastest layout on the device: 300 Voronoi grains in a disc of radius 120 px (coefficients
    U(0.5, 1.5)) plus a fibre background (orientations with the c axis within 10 deg of the sample z axis) carrying
    10 % of the coefficient sum. counts_from: an operator of the same grid whose forward projection gives the counts
    (so that every PF form solves the same data). Returns (operator, truth x_k, counts b as a device tensor,
    parameters)."""
    import torch
    from .operator import SinglePhaseForwardOperator
    p = C2
    R = uniform_fz(3.0, "hexagonal", seed=p["seed"], thin=0.6)
    cfg = scan_cfg(p["n"], p["my"], p["n_omega"], p["n_eta"])
    op = SinglePhaseForwardOperator(cfg, ti_material(), MatrixGrid(R, np.deg2rad(p["sigma_deg"])), max_gb=max_gb,
                                    normalized=True, pf_mode=pf_mode, pf_precision=pf_precision)
    K, n = op.K, p["n"]
    rng = np.random.default_rng(p["seed"])
    c = (n - 1) / 2.0
    r = p["radius"] * np.sqrt(rng.random(p["grains"]))
    t = rng.uniform(0, 2 * np.pi, p["grains"])
    sx, sy = c + r * np.cos(t), c + r * np.sin(t)
    ori = rng.choice(K, p["grains"], replace=False)
    val = rng.uniform(0.5, 1.5, p["grains"])
    yy, xx = np.mgrid[:n, :n]
    inside = ((xx - c) ** 2 + (yy - c) ** 2 <= p["radius"] ** 2).ravel()
    pix = np.nonzero(inside)[0]
    cell = np.argmin((xx.ravel()[pix, None] - sx) ** 2 + (yy.ravel()[pix, None] - sy) ** 2, axis=1)
    caxis = np.asarray(op.grid_rot_cpu)[:, :, 2]                      # crystal c axis in the sample frame
    fibre = np.nonzero(np.abs(caxis[:, 2]) >= np.cos(np.deg2rad(p["fibre_deg"])))[0]
    bg = p["fibre"] / (1 - p["fibre"]) * val.mean() / max(len(fibre), 1)
    xk = torch.zeros((n * n, op.K_stride), dtype=torch.float32, device=device)
    pt = torch.from_numpy(pix).to(device)
    kcol = torch.from_numpy(ori[cell]).to(device)           # batch 0 starts at column 0; later batches offset below
    off = np.zeros(K, np.int64)
    for bt in op.batches:
        off[bt["k_start"]:bt["k_end"]] = bt["koff"] - bt["k_start"]
    col = lambda k: torch.from_numpy(np.asarray(k) + off[np.asarray(k)]).to(device)
    xk[pt, col(ori[cell])] = torch.from_numpy(val[cell].astype(np.float32)).to(device)
    fcol = col(fibre)
    xk[pt[:, None], fcol[None, :]] += float(bg)
    del kcol
    data = torch.empty(op.data_shape, dtype=torch.float32, device=device)
    (counts_from or op).direct_k(xk.view(-1), data)
    ax = data.double()
    s = p["mean_counts"] / float(ax.mean())
    g = torch.Generator(device=device).manual_seed(p["seed"])
    b = torch.poisson(torch.clamp(ax * s, min=0), generator=g).float()
    return op, xk, b, p
