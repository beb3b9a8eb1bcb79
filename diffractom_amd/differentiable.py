"""Differentiation through the operator: the forward projection as a torch.autograd.Function whose backward is the
operator's adjoint (the HIP kernels), and the intensity layer of differentiable/data_model.py:20-64

    d[a, s, j] = f_a * t_s * r_p * u[a, s, j] + b_j          (u = A c, j = c P + p)

on torch tensors, with the reference's params0 / forward / vjp contract. The gradient of a loss L(d) with respect to
the coefficients is A^T (dL/du) (single_phase_forward_operator.py:1777-1796 adjoint); with respect to the layer's
parameters it is the reduction of (dL/dd) * u over the other axes (data_model.py:44-64).
"""
from __future__ import annotations

import ctypes

import torch

GROUPS = ("angle_scale", "trans_scale", "ring_scale", "background")


class ForwardProjection(torch.autograd.Function):
    """u = A c for coefficients c (K, Ny, Nx) float32 on the operator's device; backward: A^T g. With a subset m
    only the rows of subset m's angles are computed (the others are 0), and the backward reads only those rows."""

    @staticmethod
    def forward(ctx, coeffs, op, subset=None):
        ctx.op, ctx.subset = op, subset
        c = coeffs.detach().contiguous()
        data = torch.zeros(op.data_shape, dtype=torch.float32, device=c.device) if subset is not None else \
            torch.empty(op.data_shape, dtype=torch.float32, device=c.device)
        op.direct_cl(c, data, subset=subset)
        return data

    @staticmethod
    def backward(ctx, grad):
        op = ctx.op
        g = grad.detach().to(torch.float32).contiguous()
        out = torch.empty(op.coeff_shape, dtype=torch.float32, device=g.device)
        op.adjoint_cl(g, out, subset=ctx.subset)
        return out, None, None


def forward_project(coeffs, op, subset=None):
    """Differentiable u = A c (see ForwardProjection)."""
    return ForwardProjection.apply(coeffs, op, subset)


class IntensityLayer:
    """d = f_a t_s r_p u + b_j for data shaped (N_Omega, My, N_eta * N_rings) (data_model.py:20-64)."""

    def __init__(self, data_shape, n_rings):
        self.R, self.My, self.N_seg = (int(v) for v in data_shape)
        self.P = int(n_rings)
        self.C = self.N_seg // self.P
        if self.C * self.P != self.N_seg:
            raise ValueError(f"{self.N_seg} segments are not a multiple of {self.P} rings")

    def params0(self, device=None, dtype=torch.float64):
        """Identity parameters: unit scales, zero background."""
        kw = {"device": device, "dtype": dtype}
        return {"angle_scale": torch.ones(self.R, **kw), "trans_scale": torch.ones(self.My, **kw),
                "ring_scale": torch.ones(self.P, **kw), "background": torch.zeros(self.N_seg, **kw)}

    def _factor(self, prm):
        return (prm["angle_scale"][:, None, None, None] * prm["trans_scale"][None, :, None, None]
                * prm["ring_scale"][None, None, None, :])

    def forward(self, u, prm):
        """d (float64 unless the parameters are float32); differentiable in u and in the parameters."""
        dt = prm["angle_scale"].dtype
        u4 = u.to(dt).reshape(self.R, self.My, self.C, self.P)
        d = self._factor(prm) * u4 + prm["background"].reshape(1, 1, self.C, self.P)
        return d.reshape(self.R, self.My, self.N_seg)

    def vjp(self, u, prm, r, groups=GROUPS):
        """Gradients of <r, d> with respect to the parameter groups, and "u" = dL/du = r f t r_p (the input of the
        operator's adjoint), as data_model.py:44-64."""
        dt = prm["angle_scale"].dtype
        u4 = u.to(dt).reshape(self.R, self.My, self.C, self.P)
        r4 = r.to(dt).reshape(self.R, self.My, self.C, self.P)
        f, t, rp = prm["angle_scale"], prm["trans_scale"], prm["ring_scale"]
        ru = r4 * u4
        out = {}
        if "angle_scale" in groups:
            out["angle_scale"] = torch.einsum("ascp,s,p->a", ru, t, rp)
        if "trans_scale" in groups:
            out["trans_scale"] = torch.einsum("ascp,a,p->s", ru, f, rp)
        if "ring_scale" in groups:
            out["ring_scale"] = torch.einsum("ascp,a,s->p", ru, f, t)
        if "background" in groups:
            out["background"] = r4.sum(dim=(0, 1)).reshape(self.N_seg)
        out["u"] = (r4 * self._factor(prm)).reshape(self.R, self.My, self.N_seg)
        return out


def poisson_nll(mu, b, weights=None, eps=1e-3):
    """Poisson negative log-likelihood (deviance form, fista_poisson.py:72-109 mode 0):
    sum w [m - b + b log(b / m)], m = max(mu, 0) + eps, b log(b/m) = 0 where b = 0; float64 accumulation."""
    m = torch.clamp(mu.to(torch.float64), min=0.0) + eps
    b64 = b.to(torch.float64)
    term = m - b64 + torch.where(b64 > 0, b64 * torch.log(torch.where(b64 > 0, b64, 1.0) / m), torch.zeros_like(m))
    if weights is not None:
        term = term * weights.to(torch.float64)
    return term.sum()


class _SmoothTV(torch.autograd.Function):
    """total_variation of float32 device coefficients by the HIP kernel (value with float64 partial sums, gradient
    in the same pass)."""

    @staticmethod
    def forward(ctx, coeffs, eps):
        from . import _lib
        c = coeffs.detach().contiguous()
        grad = torch.empty_like(c)
        part = torch.empty(_lib.MAX_PARTIALS, dtype=torch.float64, device=c.device)
        n = ctypes.c_int32(0)
        K, ny, nx = c.shape
        _lib.call("dfm_smooth_tv", _lib.ptr(c), _lib.ptr(grad), K, ny, nx, float(eps), _lib.ptr(part),
                  ctypes.byref(n), _lib.stream_handle())
        ctx.save_for_backward(grad)
        return part[:n.value].sum()

    @staticmethod
    def backward(ctx, g):
        (grad,) = ctx.saved_tensors
        return grad * g.to(grad.dtype), None


def total_variation(coeffs, eps=1e-8):
    """Isotropic forward-difference TV of every orientation image (K, Ny, Nx): sum sqrt(dx^2 + dy^2 + eps), the
    last row / column differences zero. float32 device tensors: one HIP pass (value summed in float64, gradient kept
    for the backward); otherwise in float64 with torch."""
    if coeffs.is_cuda and coeffs.dtype == torch.float32 and coeffs.dim() == 3:
        return _SmoothTV.apply(coeffs, eps)
    c = coeffs.to(torch.float64)
    dx = torch.zeros_like(c)
    dy = torch.zeros_like(c)
    dx[:, :, :-1] = c[:, :, 1:] - c[:, :, :-1]
    dy[:, :-1, :] = c[:, 1:, :] - c[:, :-1, :]
    return torch.sqrt(dx * dx + dy * dy + eps).sum()


# the derivatives with respect to the model's parameters and their updates (the reference's
# diffractom.differentiable namespace)
from .gradients import (PF_GROUPS, PFGradient, ProjGradient, rotate_orientations, set_geometry,  # noqa: E402
                        set_pf_parameters, set_proj_parameters)
