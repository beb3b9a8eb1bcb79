"""Field-of-view support mask (the reference's utils/support.py)."""
import numpy as np


def fov_support_mask(Nx, Ny, n_detectors, angles=None, image_width=None, detector_width=None, detector_shift=0.0):
    """(Ny, Nx) bool: pixels whose centre projects onto the detector at every angle (parallel beam, rotation axis at
    the image centre, detector centre shifted by detector_shift). angles None: the disk of a full rotation."""
    image_width = float(Nx if image_width is None else image_width)
    detector_width = float(n_detectors if detector_width is None else detector_width)
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
