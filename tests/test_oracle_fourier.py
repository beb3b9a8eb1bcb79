"""The FP64 Fourier-slice projector (oracle/fourier.py) pinned by what the mathematics fixes: for a Gaussian blob
well inside the image and wide enough (sigma 2.5 px) that its sampling, truncation and periodization errors are below
1e-10, the projection equals the closed form: the blob's line integral sqrt(2 pi) sigma exp(-(t - t_c)^2 / (2
sigma^2)) convolved with the detector bin's hat response max(0, 1 - |s - t|) (the inverse transform of sinc^2),
computed by adaptive quadrature; also at a centre-of-rotation shift and per-angle shifts. The backward projection
is the transpose times the angle weights."""
import numpy as np
from scipy.integrate import quad

from oracle.fourier import OracleFourierRadon


def _closed_form(t_c, sig, ns, scale):
    amp = np.sqrt(2 * np.pi) * sig
    g = lambda t: amp * np.exp(-(t - t_c) ** 2 / (2 * sig ** 2))
    out = np.empty(ns)
    for s in range(ns):
        a, _ = quad(lambda t: (1 - (s - t)) * g(t), s - 1, s, epsabs=1e-13, epsrel=1e-13)
        b, _ = quad(lambda t: (1 - (t - s)) * g(t), s, s + 1, epsabs=1e-13, epsrel=1e-13)
        out[s] = scale * (a + b)
    return out


def test_gaussian_blob_projection_matches_closed_form():
    nx, ny, ns = 40, 36, 45
    angles = np.linspace(0, np.pi, 7, endpoint=False) + 0.13
    shift = np.linspace(-0.7, 0.9, 7)
    fr = OracleFourierRadon(nx, ny, angles, ns, detector_shift=0.4, shift=shift)
    c0, c1, sig = 18.3, 17.1, 2.5
    yy, xx = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
    img = np.exp(-((xx - c0) ** 2 + (yy - c1) ** 2) / (2 * sig ** 2))
    P = fr.forward(img[None])[..., 0]
    for a in range(len(angles)):
        t_c = fr.X[a] * c0 + fr.Y[a] * c1 + fr.T0[a]
        ref = _closed_form(t_c, sig, ns, fr.scale)
        assert np.abs(P[a] - ref).max() < 1e-9 * np.abs(ref).max(), a


def test_backward_is_the_weighted_transpose():
    fr = OracleFourierRadon(12, 10, np.linspace(0, np.pi, 5, endpoint=False), 15, detector_shift=0.3)
    rng = np.random.default_rng(0)
    x = rng.random((3, 10, 12))
    y = rng.random((5, 15, 3))
    w = rng.random(5)
    lhs = np.sum(fr.forward(x) * y * w[:, None, None])
    rhs = np.sum(x * fr.backward(y, w))
    assert abs(lhs / rhs - 1) < 1e-13
