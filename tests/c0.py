"""The C0 workload (Al, 5 rings, 180 omega x 103 translations x 180 eta, 99 x 99 pixels, K = 2000 random
fundamental-zone orientations, sigma 2 deg, 40 Voronoi grains, Poisson counts at a mean of 1000, SPDHG with 10
subsets for 200 steps) at full size, for the FP64 oracle: the truth is generated on the host (seeded NumPy) and the
counts are drawn from the oracle's float64 forward projection of it, so the fixture script and the GPU tests solve
the same counts. The fixture (golden/c0_oracle.npz, written by golden/make_c0_fixture.py) holds what the oracle's
SPDHG run gives; the GPU tests check the HIP path against it and against the oracle's full forward and adjoint."""
from pathlib import Path

import numpy as np

GOLDEN = Path(__file__).resolve().parent / "golden"
FIXTURE = GOLDEN / "c0_oracle.npz"
EPOCHS = 20                 # 200 subset steps with 10 subsets
DIAG = 5                    # objective every 5 epochs
N_SAMPLE = 8192             # seeded entry sample of the forward, the adjoint and the iterate


def problem():
    """(cfg, U (K, 3, 3), sigma, truth (K, 99, 99) float32, parameters, the product's (grid, material))."""
    from diffractom_amd.workloads import CONFIGS, make_grid, voronoi_truth
    p = CONFIGS["al_c0"]
    grid, mat, cfg = make_grid("al_c0")
    U = np.asarray(grid.matrices, np.float64)
    xt = voronoi_truth(len(U), p["n"], p["grains"], p["radius"], p["seed"], p["coeff"])
    return cfg, U, np.deg2rad(p["sigma_deg"]), xt, p, (grid, mat)


def oracle_operator(cfg, U, sigma):
    from diffractom_amd.workloads import AL_RINGS
    from oracle.model import OracleOperator
    orc = OracleOperator(cfg, (4.05, 4.05, 4.05, 90, 90, 90), "cubic", AL_RINGS, 0.2, U, sigma,
                         orthonormalize=False)
    orc.prefetch()
    return orc


def counts(ax, p):
    """The Poisson counts of the workload from the float64 forward projection ax (mean 1000 over all points)."""
    from diffractom_amd.workloads import poisson_counts
    b, _ = poisson_counts(ax, p["mean_counts"], p["seed"])
    return b


def sample_index(shape, seed):
    """N_SAMPLE flat indices of an array of this shape (seeded, sorted, distinct)."""
    n = int(np.prod(shape))
    return np.sort(np.random.default_rng(seed).choice(n, min(N_SAMPLE, n), replace=False))


def summary(a, seed):
    """The figures of an array the fixture keeps: the seeded entry sample, the L2 and L1 norms, the sums over the
    first axis and over all but the first axis (a checksum per orientation / per angle)."""
    a = np.asarray(a, np.float64)
    flat = a.ravel()
    idx = sample_index(a.shape, seed)
    return {"idx": idx, "val": flat[idx], "l2": np.linalg.norm(flat), "l1": np.abs(flat).sum(),
            "sum0": a.sum(axis=0).ravel(), "sum_rest": a.reshape(a.shape[0], -1).sum(axis=1)}
