"""FP64 Fourier-slice parallel-beam projector (test infrastructure: the checker of the product's projector="fft";
only tests/ import it). It restates the model the reference's FourierRadon approximates
(operators/fourier_radon.py:1-30): for angle a and detector bin s,

    P f [a, s] = scale / L sum_{j = -L/2+1}^{L/2} fhat(w_j X_a, w_j Y_a) exp(-i w_j T_a) h(w_j) exp(i w_j s),

w_j = 2 pi j / L, fhat(u, v) = sum_x f(x) exp(-i (u (x - cx) + v (y - cy))) the discrete-time Fourier transform of
the pixel samples about the reference pixel (cx, cy) = (Nx // 2, Ny // 2), T_a the detector coordinate of that
pixel, h(w) = sinc^2(w / 2 pi) the response of a detector bin and L the padded detector length (the reference's
choice: the smallest 2-3-5-7-smooth even length >= 1.25 x the span of the projections + 4). Here fhat is summed
exactly (no non-uniform FFT), so the only error is float64 rounding. Pinned by tests/test_oracle_fourier.py: a
smooth Gaussian blob's projection equals its closed form (the blob's line integral convolved with the bin's hat
response), and the backward projection is the exact transpose.
"""
import numpy as np


def smooth_length(n, even=True):
    """The smallest m >= n whose prime factors are 2, 3, 5, 7 (and even, if asked)."""
    m = int(np.ceil(n))
    while True:
        k = m
        for p in (2, 3, 5, 7):
            while k % p == 0:
                k //= p
        if k == 1 and (not even or m % 2 == 0):
            return m
        m += 1


class OracleFourierRadon:
    """P (R, Ns, Npix) of the Fourier-slice model for the ParallelRadon geometry of an operator: img (Nx, Ny),
    angles (radians), Ns detector bins of width 1, pixel width image_width / max(Nx, Ny), detector shift (bins),
    forward scale delta_x^2 / delta_s; backward = transpose times the angle weights."""

    def __init__(self, nx, ny, angles, ns, image_width=None, detector_shift=0.0, shift=None, L_factor=1.25):
        self.Nx, self.Ny, self.Ns = int(nx), int(ny), int(ns)
        self.angles = np.asarray(angles, np.float64)
        R = len(self.angles)
        dx = (float(nx) if image_width is None else float(image_width)) / max(nx, ny)
        rho = 1.0 / dx
        self.X = np.cos(self.angles - 0.5 * np.pi) / rho
        self.Y = np.sin(self.angles - 0.5 * np.pi) / rho
        sh = np.zeros(R) if shift is None else np.asarray(shift, np.float64)
        self.T0 = (ns - 1) / 2.0 - self.X * (nx - 1) / 2.0 - self.Y * (ny - 1) / 2.0 - detector_shift + sh
        self.scale = dx * dx
        self.cx, self.cy = nx // 2, ny // 2
        self.T = self.T0 + self.X * self.cx + self.Y * self.cy
        corners = np.array([[0, 0], [nx - 1, 0], [0, ny - 1], [nx - 1, ny - 1]], float)
        t = corners[:, :1] * self.X[None] + corners[:, 1:] * self.Y[None] + self.T0[None]
        span = max(t.max(), ns - 1) - min(t.min(), 0.0) + 4
        self.L = smooth_length(L_factor * span)
        yy, xx = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
        self.px, self.py = xx.ravel() - self.cx, yy.ravel() - self.cy
        self._P = None

    def matrix(self):
        """P (R, Ns, Npix) in float64 (cached)."""
        if self._P is not None:
            return self._P
        L = self.L
        j = np.arange(-L // 2 + 1, L // 2 + 1)
        w = 2 * np.pi * j / L
        h = np.sinc(w / (2 * np.pi)) ** 2
        s = np.arange(self.Ns)
        out = np.empty((len(self.angles), self.Ns, self.px.size))
        for a in range(len(self.angles)):
            # P[a, s, x] = scale / L sum_j h_j exp(i w_j (s - T_a - X (x - cx) - Y (y - cy)))
            d = s[:, None] - self.T[a] - self.X[a] * self.px[None] - self.Y[a] * self.py[None]      # (Ns, Npix)
            out[a] = self.scale / L * np.einsum("j,jsp->sp", h, np.cos(w[:, None, None] * d[None]))
        self._P = out
        return out

    def forward(self, img):
        """(R, Ns, C) projections of images (C, Ny, Nx)."""
        img = np.asarray(img, np.float64).reshape(-1, self.Ny * self.Nx)
        return np.einsum("asp,cp->asc", self.matrix(), img)

    def backward(self, sino, weights=None):
        """(C, Ny, Nx) = sum_a w_a P[a]^T sino[a] for sino (R, Ns, C); weights default 1."""
        R = len(self.angles)
        w = np.ones(R) if weights is None else np.broadcast_to(np.asarray(weights, np.float64), (R,))
        out = np.einsum("asp,asc,a->cp", self.matrix(), np.asarray(sino, np.float64), w)
        return out.reshape(-1, self.Ny, self.Nx)
