"""Small problems built twice from the same inputs: the HIP operator (diffractom_amd) and the FP64 oracle
(oracle.model). The geometries follow the reference's own tests (tests/test_subsets.py:17-28,
tests/test_differentiable.py:17-28, tests/test_spdhg_tv.py:18-51)."""
import numpy as np

from oracle.model import OracleOperator

AL3 = [[1, 1, 1], [2, 0, 0], [2, 2, 0]]


def cfg(nx=16, ny=16, my=17, n_omega=30, angle_range=(0, 180), cor=0.5, n_eta=36, wavelength=0.35):
    return {"Nx": nx, "Ny": ny, "My": my, "N_Omega": n_omega, "angle_range": list(angle_range), "cor_offset": cor,
            "N_eta": n_eta, "eta_angle_range": [0, 360], "energy": 12.398 / wavelength, "k_direction_0": [0, 0, 1],
            "p_direction_0": [1, 0, 0], "detector_direction_origin": [0, -1, 0],
            "detector_direction_positive_90": [0, 0, -1]}


def orientations(K, seed):
    from diffractom_amd.workloads import random_fz
    return random_fz(K, "cubic", seed)


def build(c=None, K=12, sigma_deg=4.0, seed=5, hkl=AL3, wavelength=0.35, oracle=True, **opkw):
    """(product operator, oracle operator, U) for cfg c."""
    from diffractom_amd import Material, MatrixGrid, SinglePhaseForwardOperator
    c = cfg() if c is None else c
    U = orientations(K, seed)
    sigma = np.deg2rad(sigma_deg)
    mat = Material.from_lattice_parameters(a=4.05, b=4.05, c=4.05, alpha=90, beta=90, gamma=90,
                                           symmetry_group="cubic", wavelength_A=wavelength, hkl_list=hkl)
    opkw.setdefault("max_gb", 0.2)
    op = SinglePhaseForwardOperator(cfg=c, material=mat, grid=MatrixGrid(U, sigma), normalized=True, **opkw)
    orc = None
    if oracle:
        quad = opkw.get("quadrature", "auto")
        if not isinstance(quad, str):
            quad = (quad.p_omega, quad.k_omega, quad.p_eta, quad.k_eta)
        orc = OracleOperator(c, (4.05, 4.05, 4.05, 90, 90, 90), "cubic", hkl, wavelength, U, sigma,
                             quadrature=quad, radon_kernel=opkw.get("radon_kernel") or "trapezoid",
                             orthonormalize=False)
    return op, orc, U


def rel_max(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def rel_l2(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))
