"""Derivatives of the forward model with respect to its continuous parameters, and updates of those parameters in
place (differentiable/gradients.py:20-258 PFGradient / ProjGradient, differentiable/update.py:10-150; the model and
its parameters as docs/differentiable_model.md defines them):

    d[a, s, j] = sum_k PF[a, k, j] S_k[a, s],  S_k = Radon(c_k),  PF from the orientations U_k (rotated by a left
    increment exp([rot_k]x)), the widths sigma_k, the Bragg angles, the eta offset and the rotation-axis tilt;
    the projection coordinate t_a(x) = X_a x + Y_a y + T0_a - cor + shift_a.

PFGradient(op)(c, r) returns the gradients of <r, A(c)> with respect to rot (K, 3), sigma (K,), eta_offset,
axis_tilt (2,) and two_theta (P,) (optionally the Gauss-Newton blocks of every orientation's rotation and width):
the projection of c (the operator's own projector), then one HIP launch per orientation batch over the PF's
candidate entries (dfm_pf_grad: the eta windows of the generated PF, float64 derivative terms), reduced in a fixed
order (dfm_pf_grad_reduce). It does not read the stored PF, so every pf_mode and precision is served.
ProjGradient(op)(c, r) returns the gradients with respect to the per-angle detector shift and the centre of
rotation: the PF adjoint of r (unscaled) against the projection of c with the footprint's derivative as weights
(the forward geometry table with d K / d t), summed per angle in float64 (dfm_angle_dots).
"""
from __future__ import annotations

import numpy as np
import torch
from scipy.spatial.transform import Rotation

from . import _lib

PF_GROUPS = ("rot", "sigma", "eta_offset", "axis_tilt", "two_theta")


def perp_basis(k):
    """Two unit vectors perpendicular to the rotation axis k: the directions the axis tilt turns it about. The same
    choice as the reference (reference.py:34-39), so that a tilt means the same rotation."""
    k = np.asarray(k, np.float64)
    k = k / np.linalg.norm(k)
    a = np.array([1.0, 0.0, 0.0]) if abs(k[0]) < 0.9 else np.array([0.0, 1.0, 0.0])
    e1 = np.cross(k, a)
    e1 /= np.linalg.norm(e1)
    return e1, np.cross(k, e1)


def _device_f32(a, shape, name, dev):
    if isinstance(a, torch.Tensor):
        if tuple(a.shape) != tuple(shape):
            raise ValueError(f"{name} must have shape {tuple(shape)}, not {tuple(a.shape)}")
        return a.to(device=dev, dtype=torch.float32).contiguous()
    a = np.asarray(a)
    if a.shape != tuple(shape):
        raise ValueError(f"{name} must have shape {tuple(shape)}, not {a.shape}")
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).to(dev)


def probed_derivatives(op, groups=PF_GROUPS):
    """Float64 derivatives of the operator's probed directions v = Rot(omega)^T u(eta, theta_p), each shaped as its
    coords (R, S, C, Se, P, 3): "eta_offset" (d v / d eta), "two_theta" (ring p's slice: d v / d (2 theta_p)),
    ("axis_tilt", i): d v / d t_i for a tilt increment t_i about e_i, e_i x v - Rot^T (e_i x u) (reference.py:
    147-166)."""
    c = op.cfg
    k0 = np.asarray(c["k_direction_0"], np.float64)
    d0 = np.asarray(c["detector_direction_origin"], np.float64)
    d90 = np.asarray(c["detector_direction_positive_90"], np.float64)
    p0 = np.asarray(c["p_direction_0"], np.float64)
    R, S, C, Se, P = op.N_Omega, op.N_Omega_subdivisions, op.N_eta, op.N_eta_subdivisions, op.N_peaks
    Rm = Rotation.from_rotvec(np.asarray(op.angles_subdivided, np.float64)[:, None] * k0).as_matrix()
    eta = np.asarray(op.eta_angles_subdivided, np.float64).reshape(C, Se)
    e = np.cos(eta)[..., None] * d0 + np.sin(eta)[..., None] * d90                        # (C, Se, 3)
    de = -np.sin(eta)[..., None] * d0 + np.cos(eta)[..., None] * d90
    th = np.asarray(op.two_theta_peaks, np.float64) / 2.0
    ct, st = np.cos(th)[None, None, :, None], np.sin(th)[None, None, :, None]
    u = ct * e[:, :, None, :] - st * p0                                                    # (C, Se, P, 3)
    rot = lambda w: np.einsum("oij,cepi->ocepj", Rm, w).reshape(R, S, C, Se, P, 3)
    out = {}
    if "eta_offset" in groups:
        out["eta_offset"] = rot(ct * de[:, :, None, :])
    if "two_theta" in groups:
        out["two_theta"] = rot(0.5 * (-st * e[:, :, None, :] - ct * p0))
    if "axis_tilt" in groups:
        v = rot(u)
        for i, ei in enumerate(perp_basis(k0)):
            out[("axis_tilt", i)] = np.cross(ei, v) - rot(np.cross(ei, u))
    return out


class PFGradient:
    """grad = PFGradient(op)(c, r): the gradients of <r, A(c)> with respect to the PF parameters at the operator's
    current parameters (gradients.py:20-189). hessian: also "hess" (K, 4, 4) = J_k^T diag(curv) J_k for the
    rotation and width of orientation k (the diagonal blocks of the Gauss-Newton matrix)."""

    def __init__(self, op, hessian=False):
        self.op = op
        self.hessian = bool(hessian)
        self.P = int(op.N_peaks)
        self._dv = {}
        self._version = getattr(op, "pf_version", 0)

    def refresh(self):
        """Forget the probed directions' derivatives (after the operator's geometry changed)."""
        self._dv = {}

    def _dv_device(self, groups):
        op = self.op
        need = [g for g in ("eta_offset", "axis_tilt", "two_theta") if g in groups and g not in self._dv]
        if need:
            host = probed_derivatives(op, need)
            for key, val in host.items():
                self._dv[key] = torch.from_numpy(np.ascontiguousarray(val)).to(op.device)
            if "axis_tilt" in need:
                self._dv["axis_tilt"] = True
        tilt = "axis_tilt" in groups
        return (self._dv["eta_offset"] if "eta_offset" in groups else None,
                self._dv[("axis_tilt", 0)] if tilt else None, self._dv[("axis_tilt", 1)] if tilt else None,
                self._dv["two_theta"] if "two_theta" in groups else None)

    def __call__(self, c, r, groups=PF_GROUPS, curv=None):
        """c: coefficients (K, Ny, Nx), r: dL / d data (N_Omega, My, N_seg), NumPy or device tensors; curv: the loss's
        curvature per data point for the Gauss-Newton blocks (hessian=True; default 1). Returns a dict with the
        requested groups ("rot" (K, 3), "sigma" (K,), "eta_offset", "axis_tilt" (2,), "two_theta" (P,)) and, with
        hessian=True, "hess" (K, 4, 4)."""
        groups = tuple(groups)
        bad = [g for g in groups if g not in PF_GROUPS]
        if bad:
            raise ValueError(f"unknown gradient groups {bad}; choose from {PF_GROUPS}")
        op = self.op
        if getattr(op, "pf_version", 0) != self._version:
            self.refresh()
            self._version = getattr(op, "pf_version", 0)
        dev = op.device
        ct = _device_f32(c, op.coeff_shape, "c", dev)
        rt = _device_f32(r, op.data_shape, "r", dev)
        cv = None
        if self.hessian:
            cv = torch.ones(op.data_shape, dtype=torch.float32, device=dev) if curv is None else \
                _device_f32(curv, op.data_shape, "curv", dev)
        geo = op._cand_geometry()
        dv_eta, dv_t0, dv_t1, dv_tth = self._dv_device(groups)
        x_k = op.to_k(ct)
        R, My, P = op.N_Omega, op.My, self.P
        nqk = 14 if self.hessian else 4
        omegas = torch.arange(R, dtype=torch.int32, device=dev)
        per_k = np.zeros((op.K, nqk))
        glob = np.zeros(3 + P)
        st = op._stream()
        qw = op.quad_weights_gpu if op.quad_weights_cpu is not None else None
        for b in op.batches:
            kb, ks = b["K_batch"], b["kstride"]
            sino = torch.empty(R * My * ks, dtype=torch.float32, device=dev)
            op._radon_forward(x_k, op.K_stride, b["koff"], ks, sino, op.geo_fwd, R, op._tables_full)
            n_wg = R * (-(-kb // 32))
            part_k = torch.empty(R * kb * nqk, dtype=torch.float64, device=dev)
            part_g = torch.empty(n_wg * (3 + P), dtype=torch.float64, device=dev)
            _lib.call("dfm_pf_grad", _lib.ptr(op.coords_gpu), _lib.ptr(op.poles_gpu), _lib.ptr(op.axis_start_gpu),
                      _lib.ptr(op.axis_count_gpu), _lib.ptr(op.inv_sigma2_gpu), _lib.ptr(op.norm_factor_gpu),
                      _lib.ptr(qw), _lib.ptr(geo["frames"]), _lib.ptr(geo["sin_th"]), _lib.ptr(geo["cos_th"]),
                      _lib.ptr(dv_eta), _lib.ptr(dv_t0), _lib.ptr(dv_t1), _lib.ptr(dv_tth), _lib.ptr(sino), ks,
                      _lib.ptr(rt), _lib.ptr(cv), _lib.ptr(omegas), R, b["k_start"], kb, My, op.N_eta, P,
                      op.N_axes, op.N_Omega_subdivisions, op.N_eta_subdivisions, float(op.pf_cut), geo["eta0"],
                      geo["deta"], geo["xlo"], geo["xhi"], _lib.ptr(part_k), _lib.ptr(part_g), st)
            out_k = torch.empty(kb * nqk, dtype=torch.float64, device=dev)
            out_g = torch.empty(3 + P, dtype=torch.float64, device=dev)
            _lib.call("dfm_pf_grad_reduce", _lib.ptr(part_k), _lib.ptr(part_g), R, kb, nqk, n_wg, 3 + P,
                      _lib.ptr(out_k), _lib.ptr(out_g), st)
            per_k[b["k_start"]:b["k_end"]] = out_k.cpu().numpy().reshape(kb, nqk)
            glob += out_g.cpu().numpy()
            del sino, part_k, part_g
        out = {}
        if "rot" in groups:
            out["rot"] = per_k[:, 0:3].copy()
        if "sigma" in groups:
            out["sigma"] = per_k[:, 3].copy()
        if "eta_offset" in groups:
            out["eta_offset"] = float(glob[0])
        if "axis_tilt" in groups:
            out["axis_tilt"] = glob[1:3].copy()
        if "two_theta" in groups:
            out["two_theta"] = glob[3:3 + P].copy()
        if self.hessian:
            iu = np.triu_indices(4)
            H = np.zeros((op.K, 4, 4))
            H[:, iu[0], iu[1]] = per_k[:, 4:14]
            H[:, iu[1], iu[0]] = per_k[:, 4:14]
            out["hess"] = H
        return out


class ProjGradient:
    """grad = ProjGradient(op)(c, r): the gradients of <r, A(c)> with respect to a detector shift of every projection
    ("shift", (N_Omega,), in bins: t_a <- t_a + shift_a) and the centre of rotation ("cor", in the operator's
    cor_offset units) (gradients.py:192-258)."""

    def __init__(self, op):
        if op.projector != "native":
            raise NotImplementedError("ProjGradient needs the native projector")
        if op._tables_full is None:
            raise NotImplementedError("ProjGradient needs the two-tap projector tables (pixels no larger than a bin)")
        self.op = op

    def __call__(self, c, r):
        op = self.op
        dev = op.device
        ct = _device_f32(c, op.coeff_shape, "c", dev)
        rt = _device_f32(r, op.data_shape, "r", dev)
        x_k = op.to_k(ct)
        tab = op._tables_full
        R, My = op.N_Omega, op.My
        st = op._stream()
        total = torch.zeros(R, dtype=torch.float64, device=dev)
        for ib, b in enumerate(op.batches):
            ks = b["kstride"]
            Z = torch.empty(R * My * ks, dtype=torch.float32, device=dev)
            D = torch.empty_like(Z)
            op._pf_adjoint(ib, rt, Z, None, alpha=1.0)                       # Z = PF^T r
            _lib.call("dfm_radon_forward_table", _lib.ptr(x_k), _lib.ptr(D), _lib.ptr(tab["row_ptr"]),
                      _lib.ptr(tab["pix"]), _lib.ptr(tab["dw"]), My, R, op.K_stride, b["koff"], ks, ks,
                      op.radon_variant["fwd"], st)                          # D = d Radon(c) / d t
            part = torch.empty(R, dtype=torch.float64, device=dev)
            _lib.call("dfm_angle_dots", _lib.ptr(Z), _lib.ptr(D), R, My, ks, b["K_batch"], _lib.ptr(part), st)
            total += part
            del Z, D
        gs = total.cpu().numpy()
        return {"shift": gs, "cor": -float(gs.sum())}


# ---------------------------------------------------------------------------------------------- updates
def rotate_orientations(rotations, increments):
    """U_k <- exp([increment_k]x) U_k (the left increment the gradients refer to). rotations (K, 3, 3)."""
    return Rotation.from_rotvec(np.asarray(increments, np.float64)).as_matrix() @ np.asarray(rotations, np.float64)


def _initial_geometry(op):
    """The geometry the operator was built with (recorded on the first update)."""
    if getattr(op, "_geo_initial", None) is None:
        op._geo_initial = {"k0": np.asarray(op.cfg["k_direction_0"], np.float64).copy(),
                           "eta": tuple(np.array(a, dtype=np.float64) for a in
                                        (op.eta_angle_range, op.eta_angles, op.eta_angles_subdivided)),
                           "two_theta": np.asarray(op.two_theta_peaks, np.float64).copy(),
                           "cor": float(op.cor_offset)}
    return op._geo_initial


def _tilt_axis(k0, axis_tilt):
    e1, e2 = perp_basis(k0)
    q = Rotation.from_rotvec(float(axis_tilt[0]) * e1 + float(axis_tilt[1]) * e2).as_matrix()
    return list(q @ np.asarray(k0, np.float64))


def set_pf_parameters(op, rotations=None, sigma=None, two_theta=None, eta_offset=0.0, axis_tilt=None):
    """Give the operator new PF parameters and rebuild what depends on them (update.py:15-56): rotations (K, 3, 3)
    orientations (crystal to sample); sigma (K,) widths (radians); two_theta (P,) Bragg angles (radians); eta_offset
    added to the azimuths (radians); axis_tilt (2,) a rotation of the current rotation axis about its two
    perpendicular directions. The quadrature rule is kept. Returns op."""
    _initial_geometry(op)
    pf = geo = False
    if rotations is not None:
        op.grid_rot_cpu = np.asarray(rotations, np.float64).reshape(op.K, 3, 3).copy()
        pf = True
    if sigma is not None:
        op.sigma_cpu = np.asarray(sigma, np.float32).reshape(op.K).copy()
        pf = True
    if two_theta is not None:
        op._ring_two_theta = np.asarray(two_theta, np.float32).reshape(op.N_peaks).copy()
        geo = True
    if eta_offset:
        d = float(eta_offset)
        op.eta_angle_range = np.asarray(op.eta_angle_range, np.float64) + d
        op.eta_angles = np.asarray(op.eta_angles, np.float64) + d
        op.eta_angles_subdivided = np.asarray(op.eta_angles_subdivided, np.float64) + d
        geo = True
    if axis_tilt is not None and np.any(np.asarray(axis_tilt) != 0):
        op.cfg = dict(op.cfg)                      # the caller's dict stays as it was
        op.cfg["k_direction_0"] = _tilt_axis(op.cfg["k_direction_0"], axis_tilt)
        geo = True
    if pf or geo:
        op.rebuild(pf=pf, geometry=geo)
    return op


def set_geometry(op, eta_offset=None, axis_tilt=None, two_theta=None, cor=None, shift=None):
    """Set the geometry absolutely, relative to the geometry the operator was built with (update.py:70-103):
    eta_offset added to the initial azimuths; axis_tilt (2,) the initial axis turned about its two perpendicular
    directions; two_theta (P,); cor; shift (N_Omega,). None keeps a parameter. Returns op."""
    g0 = _initial_geometry(op)
    geo = False
    if eta_offset is not None:
        d = float(eta_offset)
        op.eta_angle_range, op.eta_angles, op.eta_angles_subdivided = (a + d for a in g0["eta"])
        geo = True
    if axis_tilt is not None:
        op.cfg = dict(op.cfg)
        op.cfg["k_direction_0"] = _tilt_axis(g0["k0"], axis_tilt)
        geo = True
    if two_theta is not None:
        op._ring_two_theta = np.asarray(two_theta, np.float32).reshape(op.N_peaks).copy()
        geo = True
    if geo:
        op.rebuild(geometry=True)
    if cor is not None or shift is not None:
        set_proj_parameters(op, cor=cor, shift=shift)
    return op


def set_proj_parameters(op, cor=None, shift=None):
    """Give the operator a new centre of rotation (cor_offset units) and / or a detector shift of every projection
    (shift (N_Omega,), bins, added to the detector coordinate) and rebuild its projector tables (update.py:132-150;
    the PF is not affected). Returns op."""
    _initial_geometry(op)
    if cor is not None:
        op.cor_offset = float(cor)
    if shift is not None:
        op.angle_shift = np.asarray(shift, np.float64).reshape(op.N_Omega).copy()
    op.rebuild(projector=True)
    return op
