"""Fourier-slice parallel-beam projector (projector="fft"; operators/fourier_radon.py:1-311): the same geometry and
conventions as the native projector (detector coordinate t = X x + Y y + T0 in bins, forward scaled by
delta_x^2 / delta_s, backward = the transpose without that scale, weighted per angle), but the projection computed
through the Fourier slice theorem,

    P f [a, s] = scale / L sum_j fhat(w_j X_a, w_j Y_a) exp(-i w_j T_a) h(w_j) exp(i w_j s),   w_j = 2 pi j / L,

with h(w) = sinc^2(w / 2 pi) the response of a detector bin, L the padded detector length and fhat the pixels'
discrete-time Fourier transform about the pixel (Nx // 2, Ny // 2), evaluated on the polar samples by a type-2
non-uniform FFT: pixels divided by the kernel's transform psi, zero-padded onto a grid oversampled by sigma = 1.5,
a 2-D FFT, and interpolation with the "exponential of semicircle" kernel phi(z) = exp(beta (sqrt(1 - z^2) - 1)) of
W = 7 taps per dimension, beta = 0.97 pi (1 - 1 / (2 sigma)) W (the published method of FINUFFT: Barnett, Magland
and af Klinteberg, SIAM J. Sci. Comput. 41 (2019) C479-C504), about 1e-5 relative error against the exact
transform. The backward projection is the exact transpose of every step.

Device steps (csrc/fourier.hip) run in the operator's orientation-fastest layout (points x channels), the FFTs are
library calls over the leading dimensions (torch.fft, rocFFT underneath); channels go through in slices of
`max_channels`.
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp
import torch

from . import _lib


def smooth_length(n, even=True):
    """The smallest m >= n whose prime factors are 2, 3, 5, 7 (fast FFT sizes), even if asked."""
    m = int(np.ceil(n))
    while True:
        k = m
        for p in (2, 3, 5, 7):
            while k % p == 0:
                k //= p
        if k == 1 and (not even or m % 2 == 0):
            return m
        m += 1


def es_kernel(z, beta):
    """phi(z) = exp(beta (sqrt(1 - z^2) - 1)) for |z| < 1, else 0."""
    z = np.asarray(z, np.float64)
    out = np.zeros_like(z)
    inside = np.abs(z) < 1.0
    out[inside] = np.exp(beta * (np.sqrt(1.0 - z[inside] ** 2) - 1.0))
    return out


def es_transform(x, M, W, beta, n_nodes=400):
    """psi(x) = int_{-W/2}^{W/2} phi(2 k / W) cos(2 pi k x / M) dk (Gauss-Legendre quadrature)."""
    t, wq = np.polynomial.legendre.leggauss(n_nodes)
    k = 0.5 * W * t
    phi = es_kernel(2 * k / W, beta) * (0.5 * W) * wq
    return np.cos(2 * np.pi * np.outer(np.asarray(x, np.float64), k) / M) @ phi


class FourierProjector:
    """The Fourier-slice projector of an operator's geometry (its angles, detector, image, centre of rotation and
    per-angle shifts, backprojection weights)."""

    def __init__(self, op, W=7, sigma=1.5, L_factor=1.25, max_channels=2048):
        nx, ny, ns = op.Nx, op.Ny, op.My
        g = op._geo64
        X, Y, T0 = (np.asarray(g[k], np.float64) for k in ("X", "Y", "T0"))
        R = len(X)
        self.Nx, self.Ny, self.Ns, self.R = nx, ny, ns, R
        self.scale = float(op.radon_scale)
        self.max_channels = int(max_channels)
        cx, cy = nx // 2, ny // 2
        T = T0 + X * cx + Y * cy
        corners = np.array([[0, 0], [nx - 1, 0], [0, ny - 1], [nx - 1, ny - 1]], float)
        t = corners[:, :1] * X[None] + corners[:, 1:] * Y[None] + T0[None]
        span = max(t.max(), ns - 1) - min(t.min(), 0.0) + 4
        self.L = L = smooth_length(L_factor * span)
        self.H = H = L // 2 + 1
        self.W = W = int(W)
        beta = 0.97 * np.pi * (1.0 - 1.0 / (2.0 * sigma)) * W
        self.Mx, self.My = Mx, My = smooth_length(sigma * nx), smooth_length(sigma * ny)
        dev = op.device
        # pixels: deapodisation and place on the grid (offsets from the reference pixel, periodic)
        yy, xx = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
        psi_x = es_transform(np.arange(nx) - cx, Mx, W, beta)
        psi_y = es_transform(np.arange(ny) - cy, My, W, beta)
        deapod = (1.0 / (psi_y[yy] * psi_x[xx])).ravel()
        gidx = (((yy - cy) % My) * Mx + ((xx - cx) % Mx)).ravel()
        self.deapod = torch.from_numpy(deapod.astype(np.float32)).to(dev)
        self.gidx = torch.from_numpy(gidx.astype(np.int32)).to(dev)
        # polar samples (angle a, frequency j >= 0): taps, weights, bin response and phase
        j = np.arange(H)
        om = 2 * np.pi * j / L
        ku = (Mx * j[None] * X[:, None] / L).ravel()                            # grid coordinates, (R H,)
        kv = (My * j[None] * Y[:, None] / L).ravel()
        k1b = np.floor(ku - W / 2).astype(np.int64) + 1
        k2b = np.floor(kv - W / 2).astype(np.int64) + 1
        taps = np.arange(W)
        w1 = es_kernel(2 * (ku[:, None] - (k1b[:, None] + taps)) / W, beta).astype(np.float32)   # (RH, W)
        w2 = es_kernel(2 * (kv[:, None] - (k2b[:, None] + taps)) / W, beta).astype(np.float32)
        wt = (w2[:, :, None] * w1[:, None, :]).reshape(-1)                     # float32 products, [l][i]
        c1 = (k1b[:, None] + taps) % Mx
        c2 = (k2b[:, None] + taps) % My
        col = (c2[:, :, None] * Mx + c1[:, None, :]).reshape(-1)
        RH = R * H
        ptr = np.arange(0, RH * W * W + 1, W * W, dtype=np.int64)
        Ifwd = sp.csr_matrix((wt, col, ptr), shape=(RH, Mx * My))
        Itr = Ifwd.T.tocsr()                                                   # grid points x samples
        h = np.sinc(om / (2 * np.pi)) ** 2
        rowfac = (self.scale * h[None, :] * np.exp(-1j * om[None, :] * T[:, None])).astype(np.complex64)
        cj = np.full(H, 2.0)
        cj[0] = 1.0
        if L % 2 == 0:
            cj[-1] = 1.0
        wb = np.asarray(g["w"], np.float64)                                   # backprojection weight per angle
        colfac = (np.conj(rowfac.astype(np.complex128)) * (cj / L)[None, :] * (wb / self.scale)[:, None]
                  ).astype(np.complex64)
        i32 = lambda a: torch.from_numpy(np.ascontiguousarray(a, dtype=np.int32)).to(dev)
        f32 = lambda a: torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).to(dev)
        c64 = lambda a: torch.from_numpy(np.ascontiguousarray(a.ravel())).to(dev)
        self.ptr, self.col, self.w = i32(Ifwd.indptr), i32(Ifwd.indices), f32(Ifwd.data)
        self.tptr, self.smp, self.tw = i32(Itr.indptr), i32(Itr.indices), f32(Itr.data)
        self.rowfac, self.colfac = c64(rowfac), c64(colfac)
        self._apos = {}
        self.device = dev

    def _positions(self, angles):
        """apos (R,): the position of every angle in the call's list, or -1 (cached per list)."""
        key = (angles.data_ptr(), angles.numel())
        t = self._apos.get(key)
        if t is None:
            a = angles.cpu().numpy()
            p = np.full(self.R, -1, np.int32)
            p[a] = np.arange(len(a), dtype=np.int32)
            t = self._apos[key] = torch.from_numpy(p).to(self.device)
        return t

    def forward(self, img_k, img_kstride, koff, kn, sino, angles, st):
        """sino (n_ang, Ns, kn), channel-fastest = the projections of the channels koff .. koff + kn of img_k
        (Npix, img_kstride) at the angles (int32 device tensor)."""
        n_ang = angles.numel()
        npix = self.Nx * self.Ny
        sv = sino[: n_ang * self.Ns * kn].view(n_ang, self.Ns, kn)
        for c0 in range(0, kn, self.max_channels):
            c = min(self.max_channels, kn - c0)
            grid = torch.zeros(self.My * self.Mx * c, dtype=torch.complex64, device=self.device)
            _lib.call("dfm_nufft_pad", _lib.ptr(img_k), img_kstride, koff + c0, c, _lib.ptr(self.deapod),
                      _lib.ptr(self.gidx), npix, _lib.ptr(grid), st)
            G = torch.fft.fft2(grid.view(self.My, self.Mx, c), dim=(0, 1)).contiguous()
            spec = torch.empty(n_ang * self.H * c, dtype=torch.complex64, device=self.device)
            _lib.call("dfm_nufft_interp", _lib.ptr(G), _lib.ptr(self.ptr), _lib.ptr(self.col), _lib.ptr(self.w),
                      _lib.ptr(self.rowfac), _lib.ptr(angles), n_ang, self.H, c, _lib.ptr(spec), st)
            lines = torch.fft.irfft(spec.view(n_ang, self.H, c), n=self.L, dim=1)
            sv[:, :, c0:c0 + c].copy_(lines[:, : self.Ns, :])

    def backward(self, sino, angles, out_k, out_kstride, koff, kn, st, accumulate=False):
        """out_k[:, koff .. koff + kn] (+)= the weighted backprojection of sino (n_ang, Ns, kn) at the angles."""
        n_ang = angles.numel()
        npix = self.Nx * self.Ny
        apos = self._positions(angles)
        sv = sino[: n_ang * self.Ns * kn].view(n_ang, self.Ns, kn)
        n_grid = self.My * self.Mx
        for c0 in range(0, kn, self.max_channels):
            c = min(self.max_channels, kn - c0)
            pad = torch.zeros(n_ang, self.L, c, dtype=torch.float32, device=self.device)
            pad[:, : self.Ns, :].copy_(sv[:, :, c0:c0 + c])
            Z = torch.fft.rfft(pad, dim=1).contiguous()
            G = torch.empty(n_grid * c, dtype=torch.complex64, device=self.device)
            _lib.call("dfm_nufft_spread", _lib.ptr(Z), _lib.ptr(self.tptr), _lib.ptr(self.smp), _lib.ptr(self.tw),
                      _lib.ptr(self.colfac), _lib.ptr(apos), self.H, c, n_grid, _lib.ptr(G), st)
            g = torch.fft.ifft2(G.view(self.My, self.Mx, c), dim=(0, 1)).contiguous()
            _lib.call("dfm_nufft_crop", _lib.ptr(g), _lib.ptr(self.gidx), _lib.ptr(self.deapod), npix, c,
                      float(n_grid), _lib.ptr(out_k), out_kstride, koff + c0, int(accumulate), st)
