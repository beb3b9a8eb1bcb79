"""SinglePhaseForwardOperator on MI355X: data = PF @ Radon(coefficients) and its adjoint, with the reference's
Python signature (operators/single_phase_forward_operator.py:158-2206).

Device arrays are torch tensors on a ROCm device (float32, C-contiguous), where the reference takes pyopencl arrays;
NumPy in gives NumPy out. The CPU part (geometry, quadrature, pole axes) runs in float64 NumPy and is cast to float32
as in the reference; every device computation is a HIP kernel of libdiffractom_hip.so called through its C ABI.

Device layout: the orientation-fastest ("channel-fastest") images (Npix, kstride) and sinograms (angles, My,
kstride) the Radon kernels work on; coefficients in the reference's (K, Ny, Nx) layout are transposed in and out per
orientation batch (dfm_gather_k / dfm_scatter_k). The SPDHG solver keeps its arrays in the channel-fastest layout for
the whole run.
"""
from __future__ import annotations

import ctypes
import time
from typing import Any, Mapping

import numpy as np
import torch

from . import _lib
from .crystallography import Material, group_reflections_into_rings
from .quadrature import DEFAULT_TOL, QuadratureRule, setup_quadrature  # noqa: F401
from .support import fov_support_mask

KEV_ANGSTROM = 12.398       # single_phase_forward_operator.py:554
AUTO_FFT_PIXELS = 400 * 400


def _round4(n: int) -> int:
    return -(-int(n) // 4) * 4


CH_ALIGN = 64   # channel stride granule: 64 floats = 256 B, so every pixel row of a channel-fastest array and every
                # 16-quad channel block start on a cache-line boundary (two 128 B lines per block, not three)


def _round_ch(n: int) -> int:
    return -(-int(n) // CH_ALIGN) * CH_ALIGN


def _resolve_device(ctx):
    if ctx is None:
        if not torch.cuda.is_available():
            raise RuntimeError("SinglePhaseForwardOperator needs a ROCm GPU (torch.cuda.is_available() is False); "
                               "there is no CPU path")
        return torch.device("cuda", torch.cuda.current_device())
    if isinstance(ctx, torch.device):
        return ctx
    if isinstance(ctx, int):
        return torch.device("cuda", ctx)
    if isinstance(ctx, str):
        return torch.device(ctx)
    raise TypeError(f"ctx must be None, a torch.device, a device index or a device string, not {type(ctx).__name__}")


def check_device(a, shape, name, device=None):
    """A torch tensor of the given shape, float32, C-contiguous (on the operator's device); raises TypeError /
    ValueError as the reference's utils/arrays.py:13-23."""
    if not isinstance(a, torch.Tensor):
        raise TypeError(f"{name} must be a torch tensor on the GPU (or a NumPy array where accepted), "
                        f"not {type(a).__name__}")
    if a.dtype != torch.float32:
        raise TypeError(f"{name} must be float32, not {a.dtype}")
    if tuple(a.shape) != tuple(shape):
        raise ValueError(f"{name} must have shape {tuple(shape)}, not {tuple(a.shape)}")
    if not a.is_contiguous():
        raise ValueError(f"{name} must be C-contiguous")
    if device is not None and a.device != device:
        raise ValueError(f"{name} is on {a.device}, the operator on {device}")
    return a


def as_host(a, shape, name):
    """C-contiguous float32 NumPy array of the given shape: a itself if it is one (utils/arrays.py:26-32)."""
    out = np.ascontiguousarray(a, dtype=np.float32)
    if out.shape != tuple(shape):
        raise ValueError(f"{name} must have shape {tuple(shape)}, not {out.shape}")
    return out


def event_timer(device, reps):
    """fn -> median HIP-event time (ms) of reps launches after 2 warm-up launches."""
    def timeit(fn):
        ts = []
        for i in range(reps + 2):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            fn()
            e1.record()
            torch.cuda.synchronize(device)
            if i >= 2:
                ts.append(e0.elapsed_time(e1))
        return float(np.median(ts))
    return timeit


def _pick(times, tol=0.04):
    """The fastest candidate, or the first-listed one within tol of it (timing noise does not flip the choice)."""
    t_min = min(times.values())
    return next(name for name, t in times.items() if t <= t_min * (1.0 + tol))


def radon_tables(geo64, idx, nx, ny, ns, radon_scale, n_bands=1):
    """Geometry tables of the projector for the angles idx (compact order), in float64 NumPy: the forward CSR over
    (angle, bin) -> (pixel, weight), the pixel table (bin, bin, w0, w1) of the backprojection and the first bin of
    every 4 x 4 pixel tile (the C ABI's dfm_radon_forward_table / dfm_radon_backward_table). Weights are those of
    ParallelRadon (parallel_radon.py:92-124): forward scale/A K(t - s), backward (pi/N_Omega)/A K(t - s),
    K(d) = clamp(hw - |d|, 0, top); pixels touch two bins (A <= 1). band_ptr: the first entry of every (ray, band)
    for n_bands bands of ceil(ny / n_bands) image rows (dfm_radon_forward_band); a ray's entries ascend in pixel
    index, so each band is one contiguous range."""
    g = geo64
    idx = np.asarray(idx, np.int64)
    yy, xx = np.meshgrid(np.arange(ny, dtype=np.float64), np.arange(nx, dtype=np.float64), indexing="ij")
    px, py = xx.ravel(), yy.ravel()
    npix = nx * ny
    t = g["X"][idx, None] * px[None] + g["Y"][idx, None] * py[None] + g["T0"][idx, None]      # (na, npix)
    hw, top = g["hw"][idx, None], g["top"][idx, None]
    sf = np.floor(t - hw).astype(np.int64) + 1
    taps = [sf, sf + 1]
    kern = [np.clip(hw - np.abs(t - s_), 0.0, top) for s_ in taps]
    valid = [(s_ >= 0) & (s_ < ns) & (k_ > 0) for s_, k_ in zip(taps, kern)]
    fwd_scale = (radon_scale * g["inv_a"][idx])[:, None]
    bwd_scale = (g["w"][idx] * g["inv_a"][idx])[:, None]
    na = len(idx)
    rows, pix, wts, dwts = [], [], [], []
    a_idx = np.broadcast_to(np.arange(na)[:, None], t.shape)
    p_idx = np.broadcast_to(np.arange(npix)[None], t.shape)
    for s_, k_, v_ in zip(taps, kern, valid):
        rows.append((a_idx * ns + s_)[v_])
        pix.append(p_idx[v_])
        wts.append((fwd_scale * k_)[v_])
        # d K / d t of the footprint (the weight of a detector shift): -sign(t - s) where hw - |t - s| lies strictly
        # between 0 and top (differentiable/reference.py:213-223)
        q_ = hw - np.abs(t - s_)
        dwts.append((fwd_scale * np.where((q_ > 0.0) & (q_ < top), -np.sign(t - s_), 0.0))[v_])
    rows, pix, wts, dwts = np.concatenate(rows), np.concatenate(pix), np.concatenate(wts), np.concatenate(dwts)
    order = np.lexsort((pix, rows))
    row_ptr = np.zeros(na * ns + 1, np.int64)
    np.cumsum(np.bincount(rows, minlength=na * ns), out=row_ptr[1:])
    sc = [np.clip(s_, 0, ns - 1) for s_ in taps]
    wb = [np.where(v_, bwd_scale * k_, 0.0) for k_, v_ in zip(kern, valid)]
    ptab = np.empty((na, npix, 4), np.int32)
    ptab[..., 0], ptab[..., 1] = sc[0], sc[1]
    ptab[..., 2] = wb[0].astype(np.float32).view(np.int32)
    ptab[..., 3] = wb[1].astype(np.float32).view(np.int32)
    tnx, tny = -(-nx // 4), -(-ny // 4)
    tile = ((py // 4) * tnx + (px // 4)).astype(np.int64)
    lo, hi = np.minimum(sc[0], sc[1]), np.maximum(sc[0], sc[1])
    base = np.full((na, tnx * tny), np.iinfo(np.int64).max)
    top_ = np.full((na, tnx * tny), np.iinfo(np.int64).min)
    for a in range(na):
        np.minimum.at(base[a], tile, lo[a])
        np.maximum.at(top_[a], tile, hi[a])
    if np.any(top_ - base >= 8):
        raise AssertionError("a 4 x 4 pixel tile reaches more than 8 bins at some angle")
    # the same for 8 x 8 tiles (at most 16 bins; the 8 x 8 backprojection is unavailable otherwise)
    tnx8, tny8 = -(-nx // 8), -(-ny // 8)
    tile8 = ((py // 8) * tnx8 + (px // 8)).astype(np.int64)
    base8 = np.full((na, tnx8 * tny8), np.iinfo(np.int64).max)
    top8 = np.full((na, tnx8 * tny8), np.iinfo(np.int64).min)
    for a in range(na):
        np.minimum.at(base8[a], tile8, lo[a])
        np.maximum.at(top8[a], tile8, hi[a])
    tbase8 = base8.astype(np.int32) if not np.any(top8 - base8 >= 16) else np.full((1, 1), -1, np.int32)
    # line sweep: per angle the sweep class (0: image rows, |X| >= |Y|; 1: image columns) and its float64 geometry
    cls = (np.abs(g["X"][idx]) < np.abs(g["Y"][idx])).astype(np.int32)
    bh = -(-ny // n_bands)
    band_ptr = np.zeros(na * ns * n_bands + 1, np.int64)
    np.cumsum(np.bincount(rows * n_bands + (pix // nx) // bh, minlength=na * ns * n_bands), out=band_ptr[1:])
    # angle groups of one sweep class, as many angles as 1024 (and 512) rays hold (dfm_radon_forward_sweep)
    def sweep_groups(rays, per=ns):
        na_grp = max(1, rays // per)
        groups = [(c, grp) for c in (0, 1) if (cls == c).any()
                  for grp in np.array_split(np.flatnonzero(cls == c), -(-int((cls == c).sum()) // na_grp))]
        ga = np.full((max(len(groups), 1), na_grp), -1, np.int32)
        gc = np.zeros(max(len(groups), 1), np.int32)
        for i, (c, grp) in enumerate(groups):
            ga[i, :len(grp)] = grp
            gc[i] = c
        return ga, gc
    sweep_grp, sweep_gcls = sweep_groups(1024)
    # ray pairs (dfm_radon_forward_pairs_sweep): ceil(ns / 2) threads per angle, 512 per workgroup; usable when the
    # two rays' footprints on a line lie within 4 pixels for every angle: (2 hw + 1) / |coef| < 4, coef the
    # along-line coefficient (the larger of |X|, |Y|)
    coef = np.maximum(np.abs(g["X"][idx]), np.abs(g["Y"][idx]))
    pairs_ok = bool(np.all((2.0 * g["hw"][idx] + 1.0) / coef < 4.0 - 1e-9)) and -(-ns // 2) <= 512
    sweep_grp_p, sweep_gcls_p = sweep_groups(512, -(-ns // 2)) if pairs_ok else (np.full((1, 1), -1, np.int32),
                                                                         np.zeros(1, np.int32))
    return {"row_ptr": row_ptr.astype(np.int32), "pix": pix[order].astype(np.int32), "band_ptr": band_ptr.astype(np.int32),
            "tbase8": tbase8,
            "sweep_geo": np.stack([g["X"][idx], g["Y"][idx], g["T0"][idx], g["hw"][idx], g["top"][idx],
                                   radon_scale * g["inv_a"][idx]], 1).astype(np.float64),
            # the backprojection's per-angle geometry (X, Y, T0, hw, top, w / A): the grouped kernels derive the
            # pixel table's taps and weights from it on the device, by the rule of ptab above
            "bp_geo": np.stack([g["X"][idx], g["Y"][idx], g["T0"][idx], g["hw"][idx], g["top"][idx],
                                g["w"][idx] * g["inv_a"][idx]], 1).astype(np.float64),
            "sweep_grp": sweep_grp, "sweep_gcls": sweep_gcls,
            "sweep_grp_p": sweep_grp_p, "sweep_gcls_p": sweep_gcls_p, "pairs_ok": pairs_ok,
            "w": wts[order].astype(np.float32), "dw": dwts[order].astype(np.float32), "ptab": ptab,
            "tbase": base.astype(np.int32)}


class SinglePhaseForwardOperator:
    """Single-material pole-figure forward operator: data = PF @ Radon(coeffs) and its adjoint."""

    coefficient_layout = "dense"

    def __init__(self, cfg: Mapping[str, Any], material: Material, grid, max_gb: float, verbose: bool = False,
                 normalized: bool = False, ctx=None, queue=None, pf_mode: str = "auto", sparse_max_fill: float = 0.3,
                 sparse_max_fill_partial: float = 0.1, sparse_max_gb: float | None = None, projector: str = "auto",
                 radon_kernel: str | None = None, fft_max_gb: float = 1.5, reserve_coefficient_arrays: int = 3,
                 pf_cutoff_sigma: float | None = 4.0, quadrature="auto", quadrature_tol: float = DEFAULT_TOL,
                 intensity_model=None, sparse_batch_max: int = 1024, rings=None, ring_two_theta=None,
                 reflection_two_theta: bool = False, window_parallax=None, pf_precision: str = "fp32", **kwargs):
        t0 = time.perf_counter()
        self.device = _resolve_device(ctx)
        self.queue = queue                      # a torch stream, or None for the current stream at call time
        self.cfg = {**cfg, **kwargs}
        self.verbose = verbose
        self.normalized = normalized
        self.material = material
        self.grid = grid
        self.pf_cutoff_sigma = pf_cutoff_sigma
        self.pf_cut = 6.0 if pf_cutoff_sigma is None else float(pf_cutoff_sigma) ** 2 / 2
        self.reserve_coefficient_arrays = int(reserve_coefficient_arrays)
        self.sparse_batch_max = int(sparse_batch_max)
        if not normalized:
            # the ring intensities need structure factors; the materials here have no atomic basis
            # (single_phase_forward_operator.py:627-656 raises the same ValueError for such a material)
            raise ValueError("normalized=False needs the ring intensities predicted from structure factors, but the "
                             "material has no atomic basis. Use normalized=True.")
        if intensity_model is not None:
            raise NotImplementedError("intensity_model applies only with normalized=False")
        if rings is not None or ring_two_theta is not None:
            raise NotImplementedError("custom rings / ring_two_theta are not implemented")
        if reflection_two_theta:
            raise NotImplementedError("reflection_two_theta is not implemented")
        if window_parallax is not None:
            raise NotImplementedError("window_parallax is not implemented")
        if projector not in ("auto", "native", "fft", "gratopy"):
            raise ValueError(f"projector must be 'auto', 'native', 'fft' or 'gratopy', not {projector!r}")
        if projector == "gratopy":
            raise NotImplementedError("projector 'gratopy' is not implemented (the native and fft projectors are)")
        self.projector = "fft" if projector == "fft" else "native"
        if radon_kernel is None:
            radon_kernel = "trapezoid"
        if radon_kernel not in ("trapezoid", "hat"):
            raise ValueError(f"kernel must be 'hat' or 'trapezoid', not {radon_kernel!r}")
        self.radon_kernel = radon_kernel
        if pf_mode not in ("dense", "sparse", "generated", "auto"):
            raise ValueError(f"pf_mode must be 'dense', 'sparse', 'generated' or 'auto', not {pf_mode!r}")
        # PF product arithmetic on the matrix cores, FP32 accumulation: "fp32" (CSR or dense FP32 products, no
        # matrix cores), "bf16" (bf16 operands), "bf16_residual" (plus the hi/lo residual correction, ~2^-16
        # relative), "fp16" (fp16 operands with power-of-two scales per ring and per row chunk, ~2^-11 relative).
        # pf_mode "sparse" / "auto" keeps the PF as 32 x 32 blocks of its non-zeros (one copy for both directions),
        # "dense" as dense bf16 fragments (one copy per direction); "generated" evaluates the blocks of each call's
        # angles on the fly from the candidate eta windows (dfm_pfgen), nothing stored.
        if pf_precision not in ("fp32", "bf16", "bf16_residual", "fp16"):
            raise ValueError(f"pf_precision must be 'fp32', 'bf16', 'bf16_residual' or 'fp16', not {pf_precision!r}")
        if pf_precision == "fp16" and pf_mode == "dense":
            raise NotImplementedError("pf_precision 'fp16' runs on the block-sparse PF (pf_mode 'sparse' or 'auto')")
        self.pf_precision = pf_precision
        self.mfma_variant = {"fwd": 2, "adj": 2, "fwd_waves": 2048}
        self.blk_variant = {"fwd_waves": 8192, "fwd_wg": 4, "adj_wg": 4}
        self.csr_variant = {"fwd": "auto", "adj": "auto", "ksplit": "auto", "kc": 256, "tsplit": "auto", "tkc": 0,
                            "tdirect": 3}

        # ---- geometry (single_phase_forward_operator.py:376-441)
        c = self.cfg
        self.ring_reflections = group_reflections_into_rings(material)
        self.N_peaks = len(self.ring_reflections)
        self.N_eta = int(c["N_eta"])
        self.N_seg = self.N_peaks * self.N_eta
        self.Nx, self.Ny, self.My, self.N_Omega = int(c["Nx"]), int(c["Ny"]), int(c["My"]), int(c["N_Omega"])
        self.cor_offset = float(c["cor_offset"])
        vs = c.get("voxel_size")
        self.voxel_size = None if vs is None else float(vs)
        self.image_width = float(self.Nx) if vs is None else float(vs) * max(self.Nx, self.Ny)
        self.angle_range = np.array(c["angle_range"]) / 180 * np.pi
        delta = (self.angle_range[1] - self.angle_range[0]) / self.N_Omega
        self.angles = np.linspace(self.angle_range[0], self.angle_range[1], self.N_Omega, endpoint=False) + delta / 2
        self.eta_angle_range = np.array(c.get("eta_angle_range", [0, 360])) / 180 * np.pi
        eta_delta = (self.eta_angle_range[1] - self.eta_angle_range[0]) / self.N_eta
        self.eta_angles = np.linspace(self.eta_angle_range[0], self.eta_angle_range[1], self.N_eta,
                                      endpoint=False) + eta_delta / 2
        mats, sig = grid.leaf_arrays()
        self.grid_rot_cpu = np.asarray(mats, np.float64)
        self.sigma_cpu = np.asarray(sig, np.float32)
        self.K = len(self.grid_rot_cpu)
        setup_quadrature(self, quadrature, quadrature_tol, float(np.min(self.sigma_cpu)))
        self._material_parameters()
        self._detector_coordinates()
        self._pole_axes()
        self._projector_geometry()
        self._fourier_setup()
        self.pf_batch_max_gb = float(max_gb)
        self._make_batches(max_gb)
        self.omega_subsets = None
        self._sub_tables = None

        # ---- PF matrix (single_phase_forward_operator.py:500-528, 1207-1374)
        self.pf_mode = pf_mode
        self.sparse_batches = None
        self.dense_batches = None
        self.mfma_batches = None
        self.blk_batches = None
        self.gen_batches = None
        self.pf_storage = None
        self.sparse_nnz = 0
        self.sparse_fill = 0.0
        self._pf_build_args = (pf_mode, sparse_max_fill, sparse_max_gb)
        self._build_pf(pf_mode, sparse_max_fill, sparse_max_gb)
        self.pf_version = 0
        self.setup_s = time.perf_counter() - t0
        if verbose:
            print(f"PF matrix: {self.pf_mode} ({self.pf_storage}), {len(self.batches)} batches of up to "
                  f"{self.K_batch_max}; nnz {self.sparse_nnz} (fill {100 * self.sparse_fill:.3f} %); "
                  f"set up in {self.setup_s:.1f} s")

    # ================================================================== set-up
    def _material_parameters(self):
        """single_phase_forward_operator.py:627-661 (normalized: every ring intensity 1)."""
        first = [r[0] for r in self.ring_reflections]
        self.h_cpu = np.asarray(self.material.h_vecs, dtype=np.float32)[first].copy(order="C")
        self.intens_cpu = np.ones(self.N_peaks, dtype=np.float32)

    def _detector_coordinates(self):
        """Probed directions v (R, S, C, Se, P, 3) = R(omega)^T u(eta, theta_p) as float32
        (single_phase_forward_operator.py:549-622); two-theta in float32 arithmetic as the reference's."""
        from scipy.spatial.transform import Rotation
        lam = KEV_ANGSTROM / self.cfg["energy"]
        self.two_theta_peaks = (2.0 * np.arcsin(np.linalg.norm(self.h_cpu, axis=1) / (4.0 * np.pi) * lam)
                                ).astype(np.float32)
        if getattr(self, "_ring_two_theta", None) is not None:          # set_pf_parameters / set_geometry
            self.two_theta_peaks = np.asarray(self._ring_two_theta, np.float32).copy()
        Se = self.N_eta_subdivisions
        eta2 = self.eta_angles_subdivided.reshape(self.N_eta, Se)
        d0 = np.array(self.cfg["detector_direction_origin"])
        d90 = np.array(self.cfg["detector_direction_positive_90"])
        p0 = np.array(self.cfg["p_direction_0"])
        k0 = np.asarray(self.cfg["k_direction_0"])
        Rm = Rotation.from_rotvec(self.angles_subdivided[:, None] * k0).as_matrix()
        out = []
        for tt in self.two_theta_peaks:
            # the float32 two-theta, its trig in float64 (differentiable/reference.py:131-141; the reference's
            # operator takes cos/sin of the float32 half angle in float32, a 3e-8 error that 1/sigma^2 amplifies)
            half = float(tt) / 2.0
            dz = np.cos(eta2)[..., None] * d0[None, None, :] + np.sin(eta2)[..., None] * d90[None, None, :]
            dz = dz * np.cos(half) - np.sin(half) * p0
            pr = np.einsum("oij,esi->oesj", Rm, dz)
            out.append(pr.reshape(self.N_Omega, self.N_Omega_subdivisions, self.N_eta, Se, 3))
        coords = np.stack(out, axis=-2)
        self.coords_cpu = np.ascontiguousarray(coords, dtype=np.float32)
        # the PF kernel reads float64 directions (the float32 copy above is the reference's coords_cpu)
        self.coords_gpu = torch.from_numpy(np.ascontiguousarray(coords, dtype=np.float64)).to(self.device)

    def _pole_axes(self):
        """Distinct pole axes per ring and their image counts, poles[k, a] = U_k axis_a
        (single_phase_forward_operator.py:742-826)."""
        h = np.asarray(self.material.h_vecs_normed, np.float64)
        sym = np.asarray(self.material.point_group_matrices, np.float64).reshape(-1, 3, 3)
        mult = np.asarray(self.material.reflections["multiplicity"], np.float64)
        axes, counts, start = [], [], [0]
        for refl in self.ring_reflections:
            share = mult[refl] / mult[refl].sum()
            imgs = np.concatenate([sym @ h[f] for f in refl])
            wts = np.concatenate([np.full(len(sym), w) for w in share])
            first = np.argmax(np.abs(imgs) > 1e-6, axis=1)
            imgs = imgs * np.sign(imgs[np.arange(len(imgs)), first])[:, None]
            uniq, inv = np.unique(np.round(imgs, 6), axis=0, return_inverse=True)
            cnt = np.bincount(inv.ravel(), weights=wts, minlength=len(uniq))
            axes.append(uniq / np.linalg.norm(uniq, axis=1, keepdims=True))
            counts.append(cnt)
            start.append(start[-1] + len(uniq))
        self.pole_axes_cpu = np.concatenate(axes)
        self.N_axes = len(self.pole_axes_cpu)
        self.axis_start_cpu = np.asarray(start, np.int32)
        self.axis_count_cpu = np.concatenate(counts).astype(np.float32)
        if not np.allclose(np.add.reduceat(self.axis_count_cpu, self.axis_start_cpu[:-1]), len(sym)):
            raise AssertionError("pole-axis counts do not sum to the point-group order")
        poles = np.einsum("kij,aj->kai", self.grid_rot_cpu, self.pole_axes_cpu)
        dev = self.device
        self.poles_gpu = torch.from_numpy(np.ascontiguousarray(poles, np.float64)).to(dev)
        self.axis_start_gpu = torch.from_numpy(self.axis_start_cpu).to(dev)
        self.axis_count_gpu = torch.from_numpy(self.axis_count_cpu).to(dev)
        s = self.sigma_cpu.astype(np.float64)         # the float32 widths of the grid, in float64 arithmetic
        self.inv_sigma2_gpu = torch.from_numpy(1.0 / (s * s)).to(dev)
        self.norm_factor_gpu = torch.from_numpy(1.0 / (8.0 * np.pi * s * s)).to(dev)
        qw = self.quad_weights_cpu if self.quad_weights_cpu is not None else np.ones(1, np.float32)
        self.quad_weights_gpu = torch.from_numpy(np.asarray(qw, np.float32)).to(dev)

    def _projector_geometry(self):
        """ParallelRadon's geometry (parallel_radon.py:92-124): X, Y, T0 per angle, the forward scale dx^2/ds and
        the backprojection weight pi/N_Omega (single_phase_forward_operator.py:457-466), as the 8-float records of
        the C ABI."""
        n_max = max(self.Nx, self.Ny)
        dx = self.image_width / float(n_max)
        ds = 1.0
        rho = ds / dx
        ang = self.angles
        X = np.cos(ang - 0.5 * np.pi) / rho
        Y = np.sin(ang - 0.5 * np.pi) / rho
        # the centre of rotation and a detector shift of every projection (bins; set_proj_parameters)
        shift = getattr(self, "angle_shift", None)
        shift = np.zeros(self.N_Omega) if shift is None else np.asarray(shift, np.float64)
        T0 = (self.My - 1) / 2.0 - X * (self.Nx - 1) / 2.0 - Y * (self.Ny - 1) / 2.0 - self.cor_offset / ds + shift
        self.radon_scale = float(np.float32(dx * dx / ds))
        self._bp_scale = 1.0 if self.voxel_size is None else self.voxel_size ** 2
        w = np.full(self.N_Omega, np.pi / self.N_Omega * self._bp_scale)
        X32, Y32 = X.astype(np.float32), Y.astype(np.float32)
        A = np.maximum(np.abs(X32), np.abs(Y32)).astype(np.float64)
        if self.radon_kernel == "hat":
            if A.max() > 1.0 + 1e-6:
                raise ValueError("kernel='hat' needs pixels no larger than the detector bins; use 'trapezoid'")
            hw, top, inv_a = np.ones_like(A), np.ones_like(A), np.ones_like(A)
            self.n_taps = 2
        else:
            hw, top, inv_a = 0.5 * (1.0 + A), np.minimum(1.0, A), 1.0 / A
            self.n_taps = max(2, int(np.ceil(1.0 + A.max() - 1e-6)))
        if self.n_taps > 4:
            raise NotImplementedError("pixels larger than 3 detector bins (voxel_size) are not implemented")
        z = np.zeros_like(A)
        self._geo64 = dict(X=X, Y=Y, T0=T0, hw=hw, top=top, inv_a=inv_a, w=w)
        geo_f = np.stack([X, Y, T0, inv_a, hw, top, z, z], 1).astype(np.float32)      # forward: scale 1/A
        geo_b = np.stack([X, Y, T0, w * inv_a, hw, top, z, z], 1).astype(np.float32)  # backward: w/A
        self.geo_fwd_cpu, self.geo_bwd_cpu = geo_f, geo_b
        self.geo_fwd = torch.from_numpy(geo_f).to(self.device)
        self.geo_bwd = torch.from_numpy(geo_b).to(self.device)
        # the angle indices of every geometry array the projector is called with (the Fourier projector needs them)
        all_a = torch.arange(self.N_Omega, dtype=torch.int32, device=self.device)
        self._geo_angles = {id(self.geo_fwd): (self.geo_fwd, all_a), id(self.geo_bwd): (self.geo_bwd, all_a)}
        # kernel choices (identical results up to float32 rounding; set by autotune): forward "table" (geometry
        # tables, 2 taps), "band", "sweep", "psweep" (line sweeps), "arith" (footprint arithmetic, any taps);
        # backward "table" (tile and grouped forms by variant number) or "arith"
        self.radon_variant = {"fwd": 8 + 256, "bwd": 4 + 64 * 4, "fwd_kernel": "table", "bwd_kernel": "table"}
        self.tuning = None
        self.marks = None                       # a list while a profiling pass records per-launch events (mark())
        # image bands of the banded forward projection: one XCD's band of 64 channels within about 2 MB of L2
        self.radon_bands = int(min(8, max(1, -(-self.Nx * self.Ny // 8192))))
        if self.radon_bands >= 4:
            self.radon_variant["fwd_kernel"] = "band"
        self._tables_full = self._radon_tables(np.arange(self.N_Omega)) if self.n_taps == 2 else None

    def _radon_tables(self, idx):
        """Device copies of radon_tables() for the angles idx (compact order)."""
        t = radon_tables(self._geo64, idx, self.Nx, self.Ny, self.My, self.radon_scale, self.radon_bands)
        return {k: torch.from_numpy(v).to(self.device) if isinstance(v, np.ndarray) else v for k, v in t.items()}

    def _make_batches(self, max_gb, K_batch=None):
        """Orientation batches (single_phase_forward_operator.py:1026-1093): as many orientations as a dense PF batch
        (R, K_batch, C, P) of max_gb holds, at most 65532 (16-bit CSR columns)."""
        bytes_per_k = self.N_Omega * self.N_eta * self.N_peaks * 4
        if K_batch is not None:
            kb = max(1, min(int(K_batch), self.K))
        else:
            kb = max(int(float(max_gb) * 1024 ** 3 // bytes_per_k), 1)
        kb = min(kb, self.K, 65532)
        self.batches = []
        k0 = 0
        while k0 < self.K:
            k1 = min(k0 + kb, self.K)
            self.batches.append({"k_start": k0, "k_end": k1, "K_batch": k1 - k0})
            k0 = k1
        self.K_batch_max = _round_ch(max(b["K_batch"] for b in self.batches))
        # channel-fastest layout of a whole coefficient array: batch b at columns koff .. koff + round64(Kb)
        off = 0
        for b in self.batches:
            b["kstride"] = _round_ch(b["K_batch"])
            b["koff"] = off
            off += b["kstride"]
        self.K_stride = off
        for name, n in (("sinogram batch (N_Omega, My, K_batch)", self.N_Omega * self.My * self.K_batch_max),
                        ("image batch (Nx, Ny, K_batch)", self.Nx * self.Ny * self.K_batch_max)):
            if n >= 2 ** 31:
                raise ValueError(f"The {name} has {n} elements, more than 2^31 - 1; lower max_gb.")

    # ------------------------------------------------------------------ PF build
    def _stream(self):
        return _lib.stream_handle(self.queue)

    def _eval_dense_batch(self, b, out=None):
        """Dense PF of batch b, (R, kstride_b, C, P), rows beyond K_batch zero (dfm_pf_eval_dense)."""
        R, C, P = self.N_Omega, self.N_eta, self.N_peaks
        km = b["kstride"]
        if out is None:
            out = torch.empty((R, km, C, P), dtype=torch.float32, device=self.device)
        qw = self.quad_weights_cpu is not None
        _lib.call("dfm_pf_eval_dense", _lib.ptr(self.coords_gpu), _lib.ptr(self.poles_gpu),
                  _lib.ptr(self.axis_start_gpu), _lib.ptr(self.axis_count_gpu), _lib.ptr(self.inv_sigma2_gpu),
                  _lib.ptr(self.norm_factor_gpu), _lib.ptr(self.quad_weights_gpu), _lib.ptr(out), R, km, C, P,
                  self.N_axes, self.N_Omega_subdivisions, self.N_eta_subdivisions, b["k_start"], b["K_batch"],
                  float(self.pf_cut), int(qw), self._stream())
        return out

    def _csr_from_dense(self, dense, b):
        """Both CSR forms of a dense batch (count, exclusive scan, fill)."""
        R, CP = self.N_Omega, self.N_seg
        km, kb = b["kstride"], b["K_batch"]
        st = self._stream()
        dev = self.device
        cnt_f = torch.empty(R * CP, dtype=torch.int32, device=dev)
        _lib.call("dfm_pf_count_fwd", _lib.ptr(dense), _lib.ptr(cnt_f), R, km, CP, kb, st)
        cnt_a = torch.empty(R * kb, dtype=torch.int32, device=dev)
        _lib.call("dfm_pf_count_adj", _lib.ptr(dense), _lib.ptr(cnt_a), R, km, CP, kb, st)
        out = {}
        for key, cnt in (("f", cnt_f), ("a", cnt_a)):
            ptr64 = torch.zeros(cnt.numel() + 1, dtype=torch.int64, device=dev)
            torch.cumsum(cnt, 0, out=ptr64[1:])
            nnz = int(ptr64[-1].item())
            if nnz >= 2 ** 31:
                raise ValueError("Too many non-zeros in one orientation batch for int32 indices; lower max_gb.")
            out["row_ptr_" + key] = ptr64.to(torch.int32)
            out["nnz"] = nnz
        n = max(out["nnz"], 1)
        out["col_k"] = torch.empty(n, dtype=torch.int16, device=dev)
        out["val_f"] = torch.empty(n, dtype=torch.float32, device=dev)
        out["col_j"] = torch.empty(n, dtype=torch.int16, device=dev)
        out["val_a"] = torch.empty(n, dtype=torch.float32, device=dev)
        _lib.call("dfm_pf_fill_fwd", _lib.ptr(dense), _lib.ptr(out["row_ptr_f"]), _lib.ptr(out["col_k"]),
                  _lib.ptr(out["val_f"]), R, km, CP, kb, st)
        _lib.call("dfm_pf_fill_adj", _lib.ptr(dense), _lib.ptr(out["row_ptr_a"]), _lib.ptr(out["col_j"]),
                  _lib.ptr(out["val_a"]), R, km, CP, kb, st)
        return out

    def _build_pf(self, mode, max_fill, max_gb):
        R, CP = self.N_Omega, self.N_seg
        if mode == "dense":
            self.pf_mode = "dense"
            self._store_dense()
            return
        if mode == "generated" and self.pf_precision != "fp32":
            self.pf_mode = "generated"
            self._setup_generated()
            return
        # fill of the first batch decides "auto" (single_phase_forward_operator.py:1240-1265)
        b0 = self.batches[0]
        dense = self._eval_dense_batch(b0)
        nnz0 = int((dense != 0).sum().item())
        fill = nnz0 / (R * b0["K_batch"] * CP)
        if self.pf_precision != "fp32":
            if mode == "auto" and fill > max_fill and self.pf_precision != "fp16":
                del dense
                self.pf_mode = "dense"
                self._store_dense()
                return
            self.pf_mode = "sparse"
            self._store_blocks(dense)
            return
        if mode == "auto" and fill > max_fill:
            del dense
            self.pf_mode = "dense"
            if self.verbose:
                print(f"Sparse PF matrix: fill fraction {100 * fill:.2f} % > {100 * max_fill:.2f} %, dense PF path")
            self._store_dense()
            return
        if mode == "generated":
            del dense
            self.pf_mode = "generated"
            self.pf_storage = "none"
            self.sparse_nnz = int(round(fill * R * self.K * CP))
            self.sparse_fill = fill
            return
        sbs, nnz = [], 0
        for i, b in enumerate(self.batches):
            if i > 0:
                dense = self._eval_dense_batch(b, out=dense if dense.shape[1] == b["kstride"] else None)
            sb = self._csr_from_dense(dense, b)
            nnz += sb["nnz"]
            sbs.append(sb)
        del dense
        self.sparse_batches = sbs
        self.pf_mode = "sparse"
        self.pf_storage = "both"
        self.sparse_nnz = nnz
        self.sparse_fill = nnz / (R * self.K * CP)

    def _store_dense(self):
        """Dense PF batches kept on the device (the dense-PF mode): FP32 (R, kstride_b, C*P), or bf16 fragments in
        the matrix-core operand order, one copy per product direction (hi, and lo for the residual)."""
        if self.pf_precision == "fp32":
            self.dense_batches = [self._eval_dense_batch(b).reshape(self.N_Omega, b["kstride"], self.N_seg)
                                  for b in self.batches]
            self.pf_storage = "dense"
            self._count_nnz(self.dense_batches)
            return
        R, CP = self.N_Omega, self.N_seg
        res = self.pf_precision == "bf16_residual"
        st = self._stream()
        self.mfma_batches = []
        dense = None
        nnz = 0
        for b in self.batches:
            if dense is None or dense.shape[1] != b["kstride"]:
                dense = None
                dense = self._eval_dense_batch(b)
            else:
                self._eval_dense_batch(b, out=dense)
            nnz += int(torch.count_nonzero(dense).item())
            kb, ks = b["K_batch"], b["kstride"]
            mb = {"n_nt": -(-CP // 16), "n_kc": -(-kb // 32), "n_kt": -(-ks // 16), "n_jc": -(-CP // 32)}
            nf = R * mb["n_nt"] * mb["n_kc"] * 512
            na = R * mb["n_kt"] * mb["n_jc"] * 512
            for key, n in (("fwd", nf), ("adj", na)):
                mb[key + "_hi"] = torch.empty(n, dtype=torch.bfloat16, device=self.device)
                mb[key + "_lo"] = torch.empty(n, dtype=torch.bfloat16, device=self.device) if res else None
            _lib.call("dfm_mfma_pack_fwd", _lib.ptr(dense), _lib.ptr(mb["fwd_hi"]), _lib.ptr(mb["fwd_lo"]), R, ks,
                      kb, CP, mb["n_nt"], mb["n_kc"], st)
            _lib.call("dfm_mfma_pack_adj", _lib.ptr(dense), _lib.ptr(mb["adj_hi"]), _lib.ptr(mb["adj_lo"]), R, ks,
                      kb, CP, mb["n_kt"], mb["n_jc"], st)
            self.mfma_batches.append(mb)
        del dense
        self.pf_storage = "bf16 fragments" + (" + residual" if res else "")
        self.sparse_nnz = nnz
        self.sparse_fill = nnz / (R * self.K * CP)

    def _store_blocks(self, dense=None):
        """The PF of every orientation batch as 32 x 32 blocks (angle, 32 orientations, 32 segments) listing their
        non-zeros (dfm_pfblk_count / dfm_pfblk_fill): block pointers (int32), positions (uint16) and values (bf16
        hi/lo pairs, or fp16 times a power-of-two scale per ring putting the ring's largest entry in [2^13, 2^14)).
        One copy serves both product directions. dense: the first batch's dense evaluation, if already made."""
        R, CP, P = self.N_Omega, self.N_seg, self.N_peaks
        st = self._stream()
        f16 = self.pf_precision == "fp16"
        self.blk_batches = []
        nnz = 0
        for i, b in enumerate(self.batches):
            if dense is None or dense.shape[1] != b["kstride"]:
                dense = None
                dense = self._eval_dense_batch(b)
            elif i > 0:
                self._eval_dense_batch(b, out=dense)
            kb, ks = b["K_batch"], b["kstride"]
            n_kb, n_jb = -(-kb // 32), -(-CP // 32)
            cnt = torch.empty(R * n_kb * n_jb, dtype=torch.int32, device=self.device)
            _lib.call("dfm_pfblk_count", _lib.ptr(dense), _lib.ptr(cnt), R, ks, kb, CP, st)
            ptr64 = torch.zeros(cnt.numel() + 1, dtype=torch.int64, device=self.device)
            torch.cumsum(cnt, 0, out=ptr64[1:])
            n = int(ptr64[-1].item())
            if n >= 2 ** 31:
                raise ValueError("Too many non-zeros in one orientation batch for int32 indices; lower max_gb.")
            del cnt
            blk = {"ptr": ptr64.to(torch.int32), "n_kb": n_kb, "n_jb": n_jb, "nnz": n,
                   "pos": torch.empty(max(n, 1), dtype=torch.int16, device=self.device),
                   "val": torch.empty(max(n, 1), dtype=torch.int16 if f16 else torch.int32, device=self.device),
                   "ring_scale": None, "ring_inv": None}
            del ptr64
            if f16:
                # per-ring power-of-two scale: max |PF| of ring p of this batch -> [2^13, 2^14)
                mx = dense.abs().amax(dim=(0, 1, 2)).double().cpu().numpy()          # (P,)
                e = np.where(mx > 0, np.frexp(np.where(mx > 0, mx, 1.0))[1], 0)
                sc = np.ldexp(1.0, 14 - e).astype(np.float32)
                blk["ring_scale"] = torch.from_numpy(sc).to(self.device)
                blk["ring_inv"] = torch.from_numpy((1.0 / sc.astype(np.float64)).astype(np.float32)).to(self.device)
            _lib.call("dfm_pfblk_fill", _lib.ptr(dense), _lib.ptr(blk["ptr"]), _lib.ptr(blk["pos"]),
                      _lib.ptr(blk["val"]), int(f16), _lib.ptr(blk["ring_scale"]), P, R, ks, kb, CP, st)
            nnz += n
            self.blk_batches.append(blk)
        del dense
        self.pf_storage = {"bf16": "bf16 blocks", "bf16_residual": "bf16 hi/lo blocks",
                           "fp16": "fp16 blocks, ring-scaled"}[self.pf_precision]
        self.sparse_nnz = nnz
        self.sparse_fill = nnz / (R * self.K * CP)

    def _blk_mode(self, direction):
        """Mode argument of dfm_pfblk_forward / _adjoint: the precision, plus 32 for 8-wave workgroups, plus 128
        for the 4-wave form with the y-tile count compiled in (used where the count matches; same results)."""
        wg = int(self.blk_variant.get(direction + "_wg", 4))
        fixed = bool(self.blk_variant.get(direction + "_fixed", False))
        return ({"bf16": 0, "bf16_residual": 1, "fp16": 2}[self.pf_precision] + (32 if wg == 8 else 0)
                + (128 if fixed else 0))

    def _blk_split(self, n_ang, blk):
        """k split of the block-sparse forward: enough waves for blk_variant["fwd_waves"], at most 64 partial sums."""
        tiles = -(-self.My // 16)
        n_yg = -(-tiles // 8)
        waves = n_ang * n_yg * blk["n_jb"]
        return int(max(1, min(64, blk["n_kb"], -(-int(self.blk_variant.get("fwd_waves", 8192)) // waves))))

    def _blk_forward(self, ib, sino, data, subset, accumulate):
        b, blk = self.batches[ib], self.blk_batches[ib]
        n_ang, om, _, _ = self._angle_set(subset)
        ks = self._blk_split(n_ang, blk)
        part = self._mfma_part(ks * n_ang * self.My * self.N_seg) if ks > 1 else None
        _lib.call("dfm_pfblk_forward", _lib.ptr(sino), b["kstride"], b["K_batch"], _lib.ptr(blk["ptr"]),
                  _lib.ptr(blk["pos"]), _lib.ptr(blk["val"]), self._blk_mode("fwd"), _lib.ptr(blk["ring_inv"]),
                  self.N_peaks, _lib.ptr(om), n_ang, self.My, self.N_seg, _lib.ptr(data), _lib.ptr(part), ks,
                  int(accumulate), self._stream())

    def _blk_adjoint(self, ib, data, sino, subset, alpha):
        b, blk = self.batches[ib], self.blk_batches[ib]
        n_ang, om, _, _ = self._angle_set(subset)
        _lib.call("dfm_pfblk_adjoint", _lib.ptr(data), _lib.ptr(blk["ptr"]), _lib.ptr(blk["pos"]),
                  _lib.ptr(blk["val"]), self._blk_mode("adj"), _lib.ptr(blk["ring_inv"]), self.N_peaks,
                  _lib.ptr(om), n_ang, self.My, self.N_seg, b["K_batch"], _lib.ptr(sino), b["kstride"], float(alpha),
                  self._stream())

    def _setup_generated(self):
        """The generated block-sparse PF (dfm_pfgen): the geometry of the candidate windows (omega-node frames R^T d0,
        R^T d90, R^T p0; sin / cos of half the rings' two-theta; the eta nodes' place in their bin), one count pass
        over every angle per orientation batch (non-zeros per block (angle, 32 orientations, 32 segments), kept),
        and for fp16 a power-of-two scale per ring from the entries' bound 2 |G_p| max_k 1 / (8 pi sigma_k^2)
        (the pole images of a ring, both exponential terms at their peak)."""
        R, CP, P = self.N_Omega, self.N_seg, self.N_peaks
        s2 = np.asarray(self.sigma_cpu, np.float64) ** 2
        dev = self.device
        self._gen_geo = self._cand_geometry()
        cnt_axes = np.add.reduceat(self.axis_count_cpu.astype(np.float64), self.axis_start_cpu[:-1])
        bound = 2.0 * cnt_axes * float(np.max(1.0 / (8.0 * np.pi * s2)))
        sc = np.ldexp(1.0, 14 - np.ceil(np.log2(bound)).astype(int)).astype(np.float32)
        all_r = torch.arange(R, dtype=torch.int32, device=dev)
        self.gen_batches = []
        nnz = 0
        for b in self.batches:
            n_kb, n_jb = -(-b["K_batch"] // 32), -(-CP // 32)
            counts = torch.empty(R * n_kb * n_jb, dtype=torch.int32, device=dev)
            self._pfgen(0, b, all_r, R, counts=counts)
            n = int(counts.sum(dtype=torch.int64).item())
            nnz += n
            self.gen_batches.append({"counts": counts.view(R, n_kb * n_jb), "n_kb": n_kb, "n_jb": n_jb, "nnz": n,
                                     "ring_scale": torch.from_numpy(sc).to(dev),
                                     "ring_inv": torch.from_numpy((1.0 / sc.astype(np.float64)).astype(np.float32)
                                                                  ).to(dev),
                                     "groups": {}})
        self.pf_storage = {"bf16": "generated per call: bf16 blocks", "bf16_residual": "generated per call: bf16 "
                           "hi/lo blocks", "fp16": "generated per call: fp16 blocks, ring-scaled"}[self.pf_precision]
        self.sparse_nnz = nnz
        self.sparse_fill = nnz / (R * self.K * CP)
        self._gen_scratch = None
        self._gen_caps = self._gen_offs = self._gen_tags = None

    def rebuild(self, pf=False, geometry=False, projector=False):
        """Rebuild what depends on changed parameters (the differentiable model's updates): geometry (eta grids,
        rotation axis, Bragg angles: the probed directions), pf (orientations, widths: the poles), both rebuild the
        PF as it was set up (same mode, precision and budget); projector (centre of rotation, per-angle detector
        shifts: the geometry tables; the kernel choices are kept). Omega subsets are set up again; captured graphs
        and generated scratch are dropped."""
        subsets = self.omega_subsets
        if geometry:
            self._detector_coordinates()
        if pf or geometry:
            self._pole_axes()
            self._cand_geo = None
            self.sparse_batches = self.dense_batches = self.mfma_batches = self.blk_batches = self.gen_batches = None
            self._gen_scratch = None
            self._gen_caps = self._gen_offs = self._gen_tags = None
            self._build_pf(*self._pf_build_args)
        if projector:
            keep = (dict(self.radon_variant), self.tuning)
            self._projector_geometry()
            self.radon_variant, self.tuning = keep
            self._fourier_setup()
        self._scratch = None
        self.omega_subsets = None
        if subsets is not None:
            self.set_omega_subsets(subsets)
        self.scratch_generation = getattr(self, "scratch_generation", 0) + 1
        self.pf_version = getattr(self, "pf_version", 0) + 1
        return self

    def _cand_geometry(self):
        """The geometry of the PF's candidate eta windows (dfm_pfgen, dfm_pf_grad): omega-node frames R^T d0,
        R^T d90, R^T p0 (R S, 9); sin / cos of half the rings' two-theta; the eta nodes' place in their bin (bin c's
        nodes at eta0 + (c + x) deta, xlo <= x <= xhi). Device copies, cached until the geometry changes."""
        g = getattr(self, "_cand_geo", None)
        if g is not None:
            return g
        from scipy.spatial.transform import Rotation
        c = self.cfg
        e0, e1 = (float(v) for v in self.eta_angle_range)
        if e1 <= e0:
            raise NotImplementedError("the candidate eta windows need an increasing eta range")
        s2 = np.asarray(self.sigma_cpu, np.float64) ** 2
        if float(np.max(self.pf_cut * s2)) >= 1.0 - 1e-3:
            raise NotImplementedError("the candidate eta windows need cut * sigma^2 < 1 (kernels narrower than about "
                                      "20 deg)")
        k0 = np.asarray(c["k_direction_0"], np.float64)
        Rm = Rotation.from_rotvec(self.angles_subdivided[:, None] * k0).as_matrix()          # (R S, 3, 3)
        dirs = np.stack([np.asarray(c["detector_direction_origin"], np.float64),
                         np.asarray(c["detector_direction_positive_90"], np.float64),
                         np.asarray(c["p_direction_0"], np.float64)])                      # (3, 3)
        frames = np.einsum("oij,di->odj", Rm, dirs).reshape(self.N_Omega * self.N_Omega_subdivisions, 9)
        half = self.two_theta_peaks.astype(np.float64) / 2.0
        deta = (e1 - e0) / self.N_eta
        x = (np.asarray(self.eta_angles_subdivided[: self.N_eta_subdivisions], np.float64) - e0) / deta
        dev = self.device
        g = {"frames": torch.from_numpy(np.ascontiguousarray(frames)).to(dev),
             "sin_th": torch.from_numpy(np.sin(half)).to(dev), "cos_th": torch.from_numpy(np.cos(half)).to(dev),
             "eta0": e0, "deta": deta, "xlo": float(x.min()), "xhi": float(x.max())}
        self._cand_geo = g
        return g

    def _pfgen(self, mode, b, omegas, n_ang, counts=None, ptr=None, pos=None, val=None, ring_scale=None):
        g = self._gen_geo
        qw = self.quad_weights_cpu is not None
        _lib.call("dfm_pfgen", int(mode), _lib.ptr(self.coords_gpu), _lib.ptr(self.poles_gpu),
                  _lib.ptr(self.axis_start_gpu), _lib.ptr(self.axis_count_gpu), _lib.ptr(self.inv_sigma2_gpu),
                  _lib.ptr(self.norm_factor_gpu), _lib.ptr(self.quad_weights_gpu), int(qw), _lib.ptr(g["frames"]),
                  _lib.ptr(g["sin_th"]), _lib.ptr(g["cos_th"]), _lib.ptr(omegas), int(n_ang), b["k_start"],
                  b["K_batch"], self.N_eta, self.N_peaks, self.N_axes, self.N_Omega_subdivisions,
                  self.N_eta_subdivisions, float(self.pf_cut), g["eta0"], g["deta"], g["xlo"], g["xhi"],
                  _lib.ptr(counts), _lib.ptr(ptr), _lib.ptr(pos), _lib.ptr(val),
                  int(self.pf_precision == "fp16"), _lib.ptr(ring_scale), self._stream())

    def _gen_groups(self, subset):
        """The angle groups of a generated-PF call: subset m itself, or consecutive chunks of about a tenth of the
        angles for every angle (each group's PF is generated and used in turn). (key, omegas tensor, first row)."""
        if subset is not None:
            m = self._subset(subset)
            return [(("sub", self._subsets_version, m), self._sub_omegas[m], None)]
        R = self.N_Omega
        cs = max(1, -(-R // 10))
        out = []
        for r0 in range(0, R, cs):
            r1 = min(R, r0 + cs)
            key = ("rows", r0, r1)
            om = getattr(self, "_gen_chunks", {}).get(key)
            if om is None:
                om = torch.arange(r0, r1, dtype=torch.int32, device=self.device)
                self.__dict__.setdefault("_gen_chunks", {})[key] = om
            out.append((key, om, r0))
        return out

    def _gen_group(self, ib, key, om):
        """Block pointers of batch ib for the angle group om (prefix sums of the kept per-block counts; cached)."""
        gb = self.gen_batches[ib]
        grp = gb["groups"].get(key)
        if grp is None:
            c = gb["counts"].index_select(0, om.long()).reshape(-1)
            ptr64 = torch.zeros(c.numel() + 1, dtype=torch.int64, device=self.device)
            torch.cumsum(c, 0, out=ptr64[1:])
            n = int(ptr64[-1].item())
            if n >= 2 ** 31:
                raise ValueError("Too many non-zeros in one generated angle group for int32 indices; lower max_gb.")
            grp = {"ptr": ptr64.to(torch.int32), "nnz": n}
            gb["groups"][key] = grp
        return grp

    def _gen_slot(self, ib, nnz):
        """Batch ib's slot of the entry scratch (pos, val views of at least nnz entries): one slot per orientation
        batch, so that the blocks of a subset stay in place for all batches until another angle group is generated."""
        if self._gen_caps is None or self._gen_caps[ib] < nnz:
            caps = list(self._gen_caps) if self._gen_caps is not None else [1] * len(self.batches)
            caps[ib] = max(int(nnz), caps[ib])
            self._gen_alloc(caps)
        o0, o1 = int(self._gen_offs[ib]), int(self._gen_offs[ib + 1])
        return self._gen_scratch["pos"][o0:o1], self._gen_scratch["val"][o0:o1]

    def _gen_alloc(self, caps):
        """(Re)allocate the entry scratch with the slot capacities caps; forgets what the slots held."""
        f16 = self.pf_precision == "fp16"
        self._gen_scratch = None
        tot = int(sum(caps))
        self._gen_scratch = {"pos": torch.empty(tot, dtype=torch.int16, device=self.device),
                             "val": torch.empty(tot, dtype=torch.int16 if f16 else torch.int32, device=self.device)}
        self._gen_caps = list(caps)
        self._gen_offs = np.concatenate([[0], np.cumsum(caps)]).astype(np.int64)
        self._gen_tags = [None] * len(self.batches)
        self.scratch_generation = getattr(self, "scratch_generation", 0) + 1

    def _gen_prepare(self, subsets):
        """Block pointers of every (subset, batch) and slots for the largest, made when the subsets are set, so
        that the solver's steps do no host synchronisation or allocation for them."""
        caps = list(self._gen_caps) if self._gen_caps is not None else [1] * len(self.batches)
        need = list(caps)
        for m in subsets:
            key, om, _ = self._gen_groups(m)[0]
            for ib in range(len(self.batches)):
                need[ib] = max(need[ib], self._gen_group(ib, key, om)["nnz"])
        if need != caps:
            self._gen_alloc(need)

    def _gen_blocks(self, ib, key, om):
        """The PF blocks of batch ib for the angles om (block pointers cached per angle group), generated into the
        batch's scratch slot unless the slot already holds this group: a subset step's forward leaves its blocks for
        the next step's adjoint on the same subset. Under graph capture the generation is always recorded (a replay
        cannot know what the slot holds) and the slot is marked unknown. Returns the block dict the products read."""
        b, gb = self.batches[ib], self.gen_batches[ib]
        grp = self._gen_group(ib, key, om)
        pos, val = self._gen_slot(ib, grp["nnz"])
        capturing = torch.cuda.is_current_stream_capturing()
        if capturing or self._gen_tags[ib] != key:
            self._pfgen(1, b, om, om.numel(), ptr=grp["ptr"], pos=pos, val=val, ring_scale=gb["ring_scale"])
            self._gen_tags[ib] = None if capturing else key
        self.mark("pf_generate")
        return {"ptr": grp["ptr"], "pos": pos, "val": val, "n_kb": gb["n_kb"], "n_jb": gb["n_jb"],
                "ring_inv": gb["ring_inv"]}

    def _gen_forward(self, ib, sino, data, subset, accumulate):
        b = self.batches[ib]
        row = self.My * b["kstride"]
        for key, om, r0 in self._gen_groups(subset):
            blk = self._gen_blocks(ib, key, om)
            n_ang = om.numel()
            ks = self._blk_split(n_ang, blk)
            part = self._mfma_part(ks * n_ang * self.My * self.N_seg) if ks > 1 else None
            src = _lib.ptr(sino) if r0 is None else ctypes.c_void_p(sino.data_ptr() + 4 * r0 * row)
            _lib.call("dfm_pfblk_forward", src, b["kstride"], b["K_batch"], _lib.ptr(blk["ptr"]),
                      _lib.ptr(blk["pos"]), _lib.ptr(blk["val"]), self._blk_mode("fwd") + 16,
                      _lib.ptr(blk["ring_inv"]),
                      self.N_peaks, _lib.ptr(om), n_ang, self.My, self.N_seg, _lib.ptr(data), _lib.ptr(part), ks,
                      int(accumulate), self._stream())

    def _gen_adjoint(self, ib, data, sino, subset, alpha):
        b = self.batches[ib]
        row = self.My * b["kstride"]
        for key, om, r0 in self._gen_groups(subset):
            blk = self._gen_blocks(ib, key, om)
            dst = _lib.ptr(sino) if r0 is None else ctypes.c_void_p(sino.data_ptr() + 4 * r0 * row)
            _lib.call("dfm_pfblk_adjoint", _lib.ptr(data), _lib.ptr(blk["ptr"]), _lib.ptr(blk["pos"]),
                      _lib.ptr(blk["val"]), self._blk_mode("adj") + 16, _lib.ptr(blk["ring_inv"]), self.N_peaks,
                      _lib.ptr(om), om.numel(), self.My, self.N_seg, b["K_batch"], dst, b["kstride"], float(alpha),
                      self._stream())

    def _count_nnz(self, dense_batches):
        """sparse_nnz / sparse_fill of dense batches (the PF's non-zeros, as the CSR forms would hold them)."""
        self.sparse_nnz = int(sum(int(torch.count_nonzero(d).item()) for d in dense_batches))
        self.sparse_fill = self.sparse_nnz / (self.N_Omega * self.K * self.N_seg)

    def _pf_batch(self, ib):
        """The CSR pair of batch ib: stored, or evaluated now (generated mode; valid until the next call)."""
        if self.sparse_batches is not None:
            return self.sparse_batches[ib]
        b = self.batches[ib]
        return self._csr_from_dense(self._eval_dense_batch(b), b)

    # ================================================================== shapes and helpers
    @property
    def coeff_shape(self):
        return (self.K, self.Ny, self.Nx)

    @property
    def data_shape(self):
        return (self.N_Omega, self.My, self.N_seg)

    def support_mask(self):
        return fov_support_mask(self.Nx, self.Ny, self.My, angles=self.angles, image_width=self.image_width,
                                detector_width=self.My, detector_shift=self.cor_offset)

    def device_bytes(self):
        n = 0
        for t in (self.coords_gpu, self.poles_gpu):
            n += t.numel() * t.element_size()
        for sb in self.sparse_batches or []:
            n += sum(v.numel() * v.element_size() for v in sb.values() if isinstance(v, torch.Tensor))
        for d in self.dense_batches or []:
            n += d.numel() * 4
        for mb in (self.mfma_batches or []) + (self.blk_batches or []) + (self.gen_batches or []):
            n += sum(v.numel() * v.element_size() for v in mb.values() if isinstance(v, torch.Tensor))
        return n

    def free_memory(self):
        self.sparse_batches = None
        self.dense_batches = None
        self.mfma_batches = None
        self.blk_batches = None
        self.gen_batches = None
        self._gen_scratch = None
        self._gen_caps = self._gen_offs = self._gen_tags = None
        self._scratch = None

    def _scratch_buffers(self, n_angles):
        """Per-batch scratch: channel-fastest image and sinogram of the largest batch."""
        need_s = n_angles * self.My * self.K_batch_max
        sc = getattr(self, "_scratch", None)
        if sc is None or sc["sino"].numel() < need_s:
            sc = {"img": torch.empty(self.Nx * self.Ny * self.K_batch_max, dtype=torch.float32, device=self.device),
                  "sino": torch.empty(max(need_s, self.N_Omega * self.My * self.K_batch_max), dtype=torch.float32,
                                      device=self.device)}
            self._scratch = sc
            self.scratch_generation = getattr(self, "scratch_generation", 0) + 1
        return sc

    # ================================================================== subsets
    def set_omega_subsets(self, subsets):
        """Subsets of the rotation angles for direct_cl / adjoint_cl(subset=m) (single_phase_forward_operator.py:
        1541-1599): an int M (interleaved, angle i in subset i % M) or a list of angle-index arrays; returns the
        sorted index arrays; the same subsets again is a no-op. Unlike the reference, the dense PF mode supports
        subsets too (its products take the subset's angle list)."""
        R = self.N_Omega
        if np.isscalar(subsets):
            M = int(subsets)
            if not 1 <= M <= R:
                raise ValueError(f"the number of subsets must be between 1 and N_Omega = {R}")
            subsets = [np.arange(m, R, M) for m in range(M)]
        out = []
        for s in subsets:
            s = np.unique(np.asarray(s, dtype=np.int64))
            if s.size == 0 or s[0] < 0 or s[-1] >= R:
                raise ValueError(f"every subset must hold angle indices in 0 .. {R - 1}")
            out.append(s)
        old = self.omega_subsets
        if old is not None and len(old) == len(out) and all(np.array_equal(a, b) for a, b in zip(old, out)):
            return out
        self.omega_subsets = out
        self._subsets_version = getattr(self, "_subsets_version", 0) + 1
        self._sub_omegas = [torch.from_numpy(s.astype(np.int32)).to(self.device) for s in out]
        self._sub_geo_fwd = [self.geo_fwd[torch.from_numpy(s).to(self.device)].contiguous() for s in out]
        self._sub_geo_bwd = [self.geo_bwd[torch.from_numpy(s).to(self.device)].contiguous() for s in out]
        for gf, gb, om in zip(self._sub_geo_fwd, self._sub_geo_bwd, self._sub_omegas):
            self._geo_angles[id(gf)] = (gf, om)
            self._geo_angles[id(gb)] = (gb, om)
        self._sub_tables = [self._radon_tables(s) for s in out] if self.n_taps == 2 else None
        if self.gen_batches is not None:
            self._gen_prepare(range(len(out)))
        return out

    def release_omega_subsets(self):
        self.omega_subsets = None

    def _subset(self, subset):
        if self.omega_subsets is None:
            raise ValueError("no omega subsets defined: call set_omega_subsets first")
        return int(subset)

    def _angle_set(self, subset):
        """(n_angles, omegas tensor or None, forward geo, backward geo) of the full operator or subset m."""
        if subset is None:
            return self.N_Omega, None, self.geo_fwd, self.geo_bwd
        m = self._subset(subset)
        return len(self.omega_subsets[m]), self._sub_omegas[m], self._sub_geo_fwd[m], self._sub_geo_bwd[m]

    def _tables(self, subset):
        """The geometry tables of the full angle set or subset m (None when the projector needs more than 2
        taps per pixel: the arithmetic kernels then run)."""
        if subset is None:
            return self._tables_full
        return None if self._sub_tables is None else self._sub_tables[self._subset(subset)]

    def em_fused(self, subset=None):
        """True when the backprojection choice applies the EM or PDHG update itself (a grouped table variant, modes 2
        and 3 of dfm_radon_backward_table)."""
        if self.fourier is not None:
            return False
        v = int(self.radon_variant["bwd"]) if self.radon_variant["bwd"] is not None else 0
        grouped = (v // 64) % 16 > 0 or (v // 1024) % 4 == 1
        return self.radon_variant["bwd_kernel"] == "table" and grouped and self._tables(subset) is not None

    def _radon_backward(self, sino, geo, tab, n_ang, out_k, kstride, koff, kn, x=None, z=None, T=None, inv_p=0.0,
                        lam=0.0, em=False, pdhg=False):
        """img (mode 0, out_k), the SPDHG update of x, z (mode 1) or, grouped table variants only (em_fused()), the
        EM update x = x img / (T + lam) where T > 0 (em: mode 2) or the PDHG primal step x' = max(0, x - T (img +
        lam)), z = 2 x' - x (pdhg: mode 3) from the sinogram of n_ang angles."""
        st = self._stream()
        mode = 0 if x is None else (2 if em else 3 if pdhg else 1)
        if self.fourier is not None:
            if mode >= 2:
                raise ValueError("the fused EM / PDHG updates need a grouped table backprojection (em_fused())")
            q = self.queue if self.queue is not None else torch.cuda.current_stream(self.device)
            with torch.cuda.stream(q):
                ang = self._fourier_angles(geo)
                if mode == 0:
                    self.fourier.backward(sino, ang, out_k, kstride, koff, kn, st)
                    return
                npix = self.Nx * self.Ny
                dz = torch.empty(npix * kn, dtype=torch.float32, device=self.device)
                self.fourier.backward(sino, ang, dz, kn, 0, kn, st)
                off = lambda t: ctypes.c_void_p(t.data_ptr() + 4 * koff)
                _lib.call("dfm_spdhg_update_2d", off(x), off(z), off(T), _lib.ptr(dz), npix, kn, kstride, kn,
                          float(inv_p), float(lam), st)
            return
        if mode >= 2 and (tab is None or self.radon_variant["bwd_kernel"] != "table"):
            raise ValueError("the fused EM / PDHG updates need a grouped table backprojection (em_fused())")
        if tab is not None and self.radon_variant["bwd_kernel"] != "arith":
            t8 = (self.radon_variant["bwd"] // 1024) % 4 == 1       # 8x8 tiles: their first-bin table
            if t8 and tab["tbase8"].numel() == 1:
                raise ValueError("8x8 backprojection tiles reach more than 16 bins for this geometry")
            _lib.call("dfm_radon_backward_table", _lib.ptr(sino), _lib.ptr(out_k), _lib.ptr(tab["ptab"]),
                      _lib.ptr(tab["bp_geo"]), _lib.ptr(tab["tbase8" if t8 else "tbase"]), self.Nx, self.Ny, self.My, n_ang, kstride, koff,
                      kn, kn, _lib.ptr(x),
                      _lib.ptr(z), _lib.ptr(T), float(inv_p), float(lam), mode, self.radon_variant["bwd"], st)
        elif mode == 0:
            _lib.call("dfm_radon_backward", _lib.ptr(sino), _lib.ptr(out_k), _lib.ptr(geo), self.Nx, self.Ny,
                      self.My, n_ang, kstride, koff, kn, kn, self.n_taps, st)
        else:
            _lib.call("dfm_radon_backward_spdhg", _lib.ptr(sino), _lib.ptr(geo), self.Nx, self.Ny, self.My, n_ang,
                      kstride, koff, kn, kn, self.n_taps, _lib.ptr(x), _lib.ptr(z), _lib.ptr(T), float(inv_p),
                      float(lam), st)

    # ================================================================== autotuner
    def tune_projector(self, subset, img, img_kstride, koff, kn, sino, x, z, T, timeit):
        """Time the projector candidates on the given buffers (img: orientation-fastest image(s) with stride
        img_kstride, columns koff .. koff + kn; sino: scratch of the subset's angles; x, z, T: the backprojection's
        SPDHG update operands, overwritten) and keep the fastest in radon_variant; {product: (choice, times)}."""
        if self.fourier is not None:                      # one Fourier-slice projector, nothing to choose
            return {}
        n_ang, _, geo_f, geo_b = self._angle_set(subset)
        tab = self._tables(subset)
        out = {}
        rv = self.radon_variant
        saved = dict(rv)
        if tab is not None:
            cands = {}
            sl = _lib.load().dfm_radon_sweep_lines(self.Nx, self.Ny)
            if sl and self.My <= 1024:         # 1024 rays per workgroup (512 for the 64-orientation form)
                cands.update({f"sweep:{n}": ("sweep", n) for n in (8, 4, 2) if n <= sl})
                ln = max(self.Nx, self.Ny)
                for n in (2, 1):                 # 32 orientations per workgroup
                    if 2 * n * ln * 9 * 16 <= 163840:
                        cands[f"sweep:{n}+ch32"] = ("sweep", n + 16)
            if tab["pairs_ok"]:
                ln = max(self.Nx, self.Ny)
                # first the duplicated-pixel layout (conflict-free LDS reads, twice the LDS: the fastest where it
                # fits), then the plain layout, then the forms that load each staged block at its start (at most 128
                # VGPRs: two workgroups per CU)
                cands.update({f"psweep:{n}+dup": ("psweep", n + 16) for n in (2, 1)
                              if (n == 1 or ln <= 128) and 2 * n * (ln + 5) * 256 <= 163840})
                cands.update({f"psweep:{n}": ("psweep", n) for n in (2, 1, 4)
                              if 2 * n * (ln + 4) * 128 <= 163840 and (n < 4 or ln <= 128)})
                cands.update({f"psweep:{n}+np": ("psweep", n + 8) for n in (2, 1) if n == 1 or ln <= 128})
            if self.radon_bands > 1:
                cands.update({"band:8+ch128": ("band", 8 + 16), "band:8": ("band", 8), "band:4": ("band", 4),
                              "band:8+ch256": ("band", 8 + 32)})
            cands.update({"table:8+lds+ch128": ("table", 8 + 256 + 1536), "table:8+lds": ("table", 8 + 256),
                          "table:4+lds": ("table", 4 + 256)})
            times = {}
            for name, (kern, v) in cands.items():
                rv["fwd_kernel"] = kern
                if kern == "table":
                    rv["fwd"] = v
                elif kern == "band":
                    rv["fwd_band"] = v
                elif kern == "sweep":
                    rv["fwd_sweep"] = v
                elif kern == "psweep":
                    rv["fwd_psweep"] = v
                times[name] = timeit(lambda: self._radon_forward(img, img_kstride, koff, kn, sino, geo_f, n_ang, tab))
            best = _pick(times)
            kern, v = cands[best]
            saved["fwd_kernel"] = kern
            if kern == "table":
                saved["fwd"] = v
            elif kern == "band":
                saved["fwd_band"] = v
            elif kern == "sweep":
                saved["fwd_sweep"] = v
            elif kern == "psweep":
                saved["fwd_psweep"] = v
            out["radon_forward"] = (best, times)
            # the 8x8-tile forms first (the fastest at every measured configuration; the first listed within the
            # tolerance wins), then the 4x4-tile forms
            cands = {}
            if tab["tbase8"].numel() > 1:
                cands.update({f"table8:{a}x{g}": ("table", a + 64 * g + 1024)
                              for a, g in ((8, 2), (8, 4), (8, 1), (4, 2), (4, 4), (4, 1))})
            cands.update({f"table:{a}x{g}": ("table", a + 64 * g) for a, g in ((4, 4), (4, 2), (8, 4), (8, 2), (8, 8))})
            times = {}
            for name, (kern, v) in cands.items():
                rv["bwd_kernel"] = kern
                if v is not None:
                    rv["bwd"] = v
                times[name] = timeit(lambda: self._radon_backward(sino, geo_b, tab, n_ang, None, img_kstride, koff, kn,
                                                                  x=x, z=z, T=T, inv_p=1.0))
            best = _pick(times)
            kern, v = cands[best]
            saved["bwd_kernel"] = kern
            if v is not None:
                saved["bwd"] = v
            out["radon_backward_spdhg"] = (best, times)
        self.radon_variant.update(saved)
        return out

    def autotune(self, subset=0, reps=5):
        """Pick each product's fastest kernel variant by timing it on this operator's shapes (HIP events, median of
        reps after 2 warm-up launches), as the reference times its tiled and plain projectors
        (parallel_radon.py:240-281). Every candidate computes the same values (up to float32 summation order).
        Candidates: forward projection (line sweeps with one or two rays per thread, pixel bands, geometry table
        with 4 or 8 entries fetched ahead, staged in LDS or not, capped occupancy), backprojection with the SPDHG
        update (4x4 and 8x8 tiles, 4, 8 or 16 angles staged per pass, 1-8 channel blocks per workgroup), the PF
        products (block workgroup size and k split, CSR forms).
        Returns {product: (choice, {candidate: ms})} and keeps it in self.tuning."""
        if self.omega_subsets is None:
            self.set_omega_subsets(10)
        n_ang, _, geo_f, geo_b = self._angle_set(subset)
        tab = self._tables(subset)
        sc = self._scratch_buffers(n_ang)
        b = self.batches[0]
        dev = self.device
        npix = self.Nx * self.Ny
        g = torch.Generator(device=dev)
        g.manual_seed(0)
        x = torch.rand(npix * self.K_stride, generator=g, device=dev)
        z = torch.zeros_like(x)
        T = torch.zeros_like(x)

        timeit = event_timer(dev, reps)

        out = self.tune_projector(subset, x, self.K_stride, b["koff"], b["kstride"], sc["sino"], x.clone(), z, T,
                                  timeit)
        self._autotune_pf(subset, sc, g, timeit, out)
        del x, z, T
        self.tuning = out
        return out

    def _autotune_pf(self, subset, sc, g, timeit, out):
        """The PF products' part of autotune: times every candidate on random data rows and the sinogram scratch of
        batch 0, keeps the fastest and records {product: (choice, {candidate: ms})} in out."""
        dev = self.device
        if self.sparse_batches is not None:
            data = torch.rand(self.data_shape, generator=g, device=dev)
            times = {}
            cands = {}
            if _lib.load().dfm_spmm_t_rows_per_block(self.My) > 0:
                cands.update({f"t:split:{k}": dict(fwd="t", tsplit=k) for k in (16, 8, 4, 32)})
            cands.update({f"rows:ksplit:{k},kc:{c}": dict(fwd="rows", ksplit=k, kc=c) for k in (16, 4)
                          for c in (256, 768)})
            for name, v in cands.items():
                self.csr_variant.update(v)
                times[name] = timeit(lambda: self._pf_forward(0, sc["sino"], data, subset, False))
            best = _pick(times)
            self.csr_variant.update(cands[best])
            out["pf_forward"] = (best, times)
            times = {}
            for name in (("t", "rows") if _lib.load().dfm_spmm_t_rows_per_block(self.My) > 0 else ("rows",)):
                self.csr_variant["adj"] = name
                times[name] = timeit(lambda: self._pf_adjoint(0, data, sc["sino"], subset))
            best = _pick(times)
            self.csr_variant["adj"] = best
            out["pf_adjoint"] = (best, times)
        if self.mfma_batches is not None:
            data = torch.rand(self.data_shape, generator=g, device=dev)
            fwd = lambda: self._dense_forward(0, sc["sino"], data, subset, False)
            times, cands = {}, {}
            for v in (2, 4):
                for w in (2048, 4096, 8192, 16384):
                    cands[f"tiles_per_wave:{v}" + ("" if w == 2048 else f"+waves:{w}")] = (v, w)
            for name, (v, w) in cands.items():
                self.mfma_variant.update(fwd=v, fwd_waves=w)
                times[name] = timeit(fwd)
            best = _pick(times)
            self.mfma_variant.update(fwd=cands[best][0], fwd_waves=cands[best][1])
            out["pf_forward"] = (best, times)
            times = {}
            for v in (2, 4):
                self.mfma_variant["adj"] = v
                times[f"tiles_per_wave:{v}"] = timeit(lambda: self._dense_adjoint(0, data, sc["sino"], subset, 1.0))
            best = _pick(times)
            self.mfma_variant["adj"] = int(best.split(":")[1])
            out["pf_adjoint"] = (best, times)
        if self.blk_batches is not None or self.gen_batches is not None:
            # block-sparse products (stored or generated; a generated subset's blocks are made by the first call and
            # reused by the timed ones): 4- or 8-wave workgroups, and the forward's k split
            data = torch.rand(self.data_shape, generator=g, device=dev)
            gen = self.blk_batches is None
            fwd = (lambda: self._gen_forward(0, sc["sino"], data, subset, False)) if gen else \
                (lambda: self._blk_forward(0, sc["sino"], data, subset, False))
            adj = (lambda: self._gen_adjoint(0, data, sc["sino"], subset, 1.0)) if gen else \
                (lambda: self._blk_adjoint(0, data, sc["sino"], subset, 1.0))
            times, cands = {}, {}
            for w in (4096, 8192, 16384):             # the y-tile count compiled in (4-wave form)
                cands[f"wg:4+waves:{w}+mt"] = (4, w, True)
            for wg in (4, 8):
                for w in (2048, 4096, 8192, 16384, 32768):
                    cands[f"wg:{wg}+waves:{w}"] = (wg, w, False)
            for name, (wg, w, fx) in cands.items():
                self.blk_variant.update(fwd_wg=wg, fwd_waves=w, fwd_fixed=fx)
                times[name] = timeit(fwd)
            best = _pick(times)
            self.blk_variant.update(fwd_wg=cands[best][0], fwd_waves=cands[best][1], fwd_fixed=cands[best][2])
            out["pf_forward"] = (best, times)
            times, cands = {}, {"wg:4+mt": (4, True), "wg:4": (4, False), "wg:8": (8, False)}
            for name, (wg, fx) in cands.items():
                self.blk_variant.update(adj_wg=wg, adj_fixed=fx)
                times[name] = timeit(adj)
            best = _pick(times)
            self.blk_variant.update(adj_wg=cands[best][0], adj_fixed=cands[best][1])
            out["pf_adjoint"] = (best, times)

    def zero_subset_rows(self, data, subset):
        m = self._subset(subset)
        _lib.call("dfm_zero_rows", _lib.ptr(data), _lib.ptr(self._sub_omegas[m]), len(self.omega_subsets[m]),
                  self.My * self.N_seg, self._stream())

    # ================================================================== forward
    def _fourier_setup(self):
        """The Fourier-slice projector of the current geometry (projector="fft"), else None."""
        self.fourier = None
        if self.projector == "fft":
            from .fourier import FourierProjector
            self.fourier = FourierProjector(self)

    def _fourier_angles(self, geo):
        entry = self._geo_angles.get(id(geo))
        if entry is None or entry[0] is not geo:
            raise ValueError("projector='fft': the geometry array is not one of the operator's angle sets")
        return entry[1]

    def _radon_forward(self, img, img_kstride, koff, kn, sino, geo, n_angles, tab=None):
        if self.fourier is not None:
            q = self.queue if self.queue is not None else torch.cuda.current_stream(self.device)
            with torch.cuda.stream(q):
                self.fourier.forward(img, img_kstride, koff, kn, sino, self._fourier_angles(geo), self._stream())
            return
        if tab is not None and self.radon_variant["fwd_kernel"] == "psweep" and tab["pairs_ok"]:
            g = tab["sweep_grp_p"]
            _lib.call("dfm_radon_forward_pairs_sweep", _lib.ptr(img), _lib.ptr(sino), _lib.ptr(tab["sweep_geo"]),
                      _lib.ptr(g), _lib.ptr(tab["sweep_gcls_p"]), g.shape[0], g.shape[1], self.My, self.Nx, self.Ny,
                      img_kstride, koff, kn, kn, int(self.radon_variant.get("fwd_psweep", 2)), self._stream())
            return
        if tab is not None and self.radon_variant["fwd_kernel"] in ("sweep", "psweep"):
            lines = self.radon_variant.get("fwd_sweep", 0) or _lib.load().dfm_radon_sweep_lines(self.Nx, self.Ny)
            g = tab["sweep_grp"]
            _lib.call("dfm_radon_forward_sweep", _lib.ptr(img), _lib.ptr(sino), _lib.ptr(tab["sweep_geo"]),
                      _lib.ptr(g), _lib.ptr(tab["sweep_gcls"]), g.shape[0], g.shape[1], self.My, self.Nx,
                      self.Ny, img_kstride, koff, kn, kn, int(lines), self._stream())
            return
        if tab is not None and self.radon_variant["fwd_kernel"] == "band":
            nb = self.radon_bands
            part = self._t_buf(nb * n_angles * self.My * kn)[0]
            _lib.call("dfm_radon_forward_band", _lib.ptr(img), _lib.ptr(sino), _lib.ptr(part),
                      _lib.ptr(tab["band_ptr"]), _lib.ptr(tab["pix"]), _lib.ptr(tab["w"]), self.My, n_angles, nb,
                      img_kstride, koff, kn, kn, self.radon_variant.get("fwd_band", 8), self._stream())
            return
        if tab is not None and self.radon_variant["fwd_kernel"] != "arith":
            _lib.call("dfm_radon_forward_table", _lib.ptr(img), _lib.ptr(sino), _lib.ptr(tab["row_ptr"]),
                      _lib.ptr(tab["pix"]), _lib.ptr(tab["w"]), self.My, n_angles, img_kstride, koff, kn, kn,
                      self.radon_variant["fwd"], self._stream())
            return
        _lib.call("dfm_radon_forward", _lib.ptr(img), _lib.ptr(sino), _lib.ptr(geo), self.Nx, self.Ny, self.My,
                  n_angles, img_kstride, koff, kn, kn, self.radon_scale, self.n_taps, self._stream())

    def _pf_forward(self, ib, sino, data, subset, accumulate):
        """data rows (+)= PF_b sino for batch ib (sparse CSR product or dense GEMM)."""
        b = self.batches[ib]
        n_ang, om, _, _ = self._angle_set(subset)
        if self.blk_batches is not None:
            self._blk_forward(ib, sino, data, subset, accumulate)
            return
        if self.gen_batches is not None:
            self._gen_forward(ib, sino, data, subset, accumulate)
            return
        if self.pf_mode == "dense":
            self._dense_forward(ib, sino, data, subset, accumulate)
            return
        sb = self._pf_batch(ib)
        if self._csr_kind("fwd") == "t":
            self._spmm_t_forward(b, sb, om, n_ang, sino, data, accumulate)
            return
        ks = self._csr_split(n_ang, b["K_batch"])
        part = self._mfma_part(ks * n_ang * self.My * self.N_seg) if ks > 1 else None
        _lib.call("dfm_spmm_forward_split", _lib.ptr(sino), _lib.ptr(sb["row_ptr_f"]), _lib.ptr(sb["col_k"]),
                  _lib.ptr(sb["val_f"]), _lib.ptr(data), _lib.ptr(om), n_ang, self.My, self.N_seg, b["kstride"],
                  b["K_batch"], 0, int(accumulate), _lib.ptr(part), ks, int(self.csr_variant["kc"]),
                  self._stream())

    def _pf_adjoint(self, ib, data, sino, subset, alpha=None):
        """sino = alpha PF_b^T data rows for batch ib (alpha: the adjoint's R / pi by default)."""
        b = self.batches[ib]
        n_ang, om, _, _ = self._angle_set(subset)
        alpha = float(self.N_Omega / np.pi) if alpha is None else float(alpha)
        if self.blk_batches is not None:
            self._blk_adjoint(ib, data, sino, subset, alpha)
            return
        if self.gen_batches is not None:
            self._gen_adjoint(ib, data, sino, subset, alpha)
            return
        if self.pf_mode == "dense":
            self._dense_adjoint(ib, data, sino, subset, alpha)
            return
        sb = self._pf_batch(ib)
        if self._csr_kind("adj") == "t":
            self._spmm_t_adjoint(b, sb, om, n_ang, data, sino, alpha)
            return
        _lib.call("dfm_spmm_adjoint", _lib.ptr(data), _lib.ptr(sb["row_ptr_a"]), _lib.ptr(sb["col_j"]),
                  _lib.ptr(sb["val_a"]), _lib.ptr(sino), _lib.ptr(om), n_ang, self.My, self.N_seg, b["kstride"],
                  b["K_batch"], alpha, 0, self._stream())

    # ------------------------------------------------------------------ CSR products on transposed operands
    def _csr_kind(self, direction):
        """"t" (wavefront-segmented kernel on transposed operands) or "rows" (a thread per CSR row)."""
        v = self.csr_variant.get(direction, "auto")
        if v == "auto":
            return "t" if _lib.load().dfm_spmm_t_rows_per_block(self.My) > 0 else "rows"
        return v

    @property
    def _ys(self):
        return -(-self.My // 4) * 4

    def _t_buf(self, *sizes):
        """Views of one reusable workspace, each starting 256-byte aligned."""
        offs, n = [], 0
        for z in sizes:
            offs.append(n)
            n += -(-int(z) // 64) * 64
        p = getattr(self, "_t_buf_t", None)
        if p is None or p.numel() < n:
            p = torch.empty(n, dtype=torch.float32, device=self.device)
            self._t_buf_t = p
            self.scratch_generation = getattr(self, "scratch_generation", 0) + 1
        return [p[o:o + int(z)] for o, z in zip(offs, sizes)]

    def _tsplit(self, n_ang, kb):
        """Orientation split of the transposed forward: enough workgroups for 4 per CU on 256 CUs."""
        v = self.csr_variant.get("tsplit", "auto")
        if v != "auto":
            return int(max(1, min(int(v), max(kb, 1))))
        rpb = _lib.load().dfm_spmm_t_rows_per_block(self.My)
        wgs = -(-self.N_seg // rpb) * n_ang
        return int(max(1, min(64, -(-1024 // wgs), max(kb // 128, 1))))

    def _spmm_t_forward(self, b, sb, om, n_ang, sino, data, accumulate):
        """sino (n_ang, my, kstride) -> split partials (ns, n_ang, cp, ys) -> data rows; the sinogram is transposed
        while the kernel stages it in LDS (bit 2 of csr_variant["tdirect"]) or first into sT (n_ang, kb, ys)."""
        my, ys, cp, kb = self.My, self._ys, self.N_seg, b["K_batch"]
        ns = self._tsplit(n_ang, kb)
        direct = int(self.csr_variant.get("tdirect", 3)) & 2
        st = self._stream()
        if direct:
            (part,) = self._t_buf(ns * n_ang * cp * ys)
            src, in_ld = sino, b["kstride"]
        else:
            src, part = self._t_buf(n_ang * kb * ys, ns * n_ang * cp * ys)
            in_ld = 0
            _lib.call("dfm_transpose_sum", _lib.ptr(sino), _lib.ptr(src), None, 0, n_ang, my, kb, b["kstride"],
                      my * b["kstride"], ys, kb * ys, 1, 0, 0, st)
        _lib.call("dfm_spmm_t_direct", _lib.ptr(src), in_ld, ys, _lib.ptr(sb["row_ptr_f"]), _lib.ptr(sb["col_k"]),
                  _lib.ptr(sb["val_f"]), _lib.ptr(part), 0, ys, _lib.ptr(om), n_ang, my, kb, cp, cp, 0, 1.0, ns,
                  int(self.csr_variant.get("tkc", 0)), st)
        _lib.call("dfm_transpose_sum", _lib.ptr(part), _lib.ptr(data), _lib.ptr(om), 1, n_ang, cp, my, ys, cp * ys,
                  cp, my * cp, ns, n_ang * cp * ys, int(accumulate), st)

    def _spmm_t_adjoint(self, b, sb, om, n_ang, data, sino, alpha):
        """data rows -> dT (n_ang, cp, ys) -> sino (n_ang, my, kstride), written through the kernel's LDS transpose
        (bit 1 of csr_variant["tdirect"]) or as sT (n_ang, kstride, ys) and transposed."""
        my, ys, cp, ks = self.My, self._ys, self.N_seg, b["kstride"]
        direct = int(self.csr_variant.get("tdirect", 3)) & 1
        st = self._stream()
        if direct:
            (dT,) = self._t_buf(n_ang * cp * ys)
        else:
            dT, sT = self._t_buf(n_ang * cp * ys, n_ang * ks * ys)
        _lib.call("dfm_transpose_sum", _lib.ptr(data), _lib.ptr(dT), _lib.ptr(om), 0, n_ang, my, cp, cp, my * cp, ys,
                  cp * ys, 1, 0, 0, st)
        if direct:
            _lib.call("dfm_spmm_t_direct", _lib.ptr(dT), 0, ys, _lib.ptr(sb["row_ptr_a"]), _lib.ptr(sb["col_j"]),
                      _lib.ptr(sb["val_a"]), _lib.ptr(sino), ks, 0, _lib.ptr(om), n_ang, my, cp, b["K_batch"], ks, 0,
                      alpha, 1, int(self.csr_variant.get("tkc", 0)), st)
            return
        _lib.call("dfm_spmm_t", _lib.ptr(dT), _lib.ptr(sb["row_ptr_a"]), _lib.ptr(sb["col_j"]), _lib.ptr(sb["val_a"]),
                  _lib.ptr(sT), _lib.ptr(om), n_ang, my, ys, cp, b["K_batch"], ks, 0, alpha, 1,
                  int(self.csr_variant.get("tkc", 0)), st)
        _lib.call("dfm_transpose_sum", _lib.ptr(sT), _lib.ptr(sino), None, 0, n_ang, ks, my, ys, ks * ys, ks,
                  my * ks, 1, 0, 0, st)

    def _mfma_split(self, n_ang, mb):
        """k split of the matrix-core forward: enough waves for mfma_variant["fwd_waves"] (2048: 8 per CU on 256
        CUs), at most 32 partial sums."""
        tiles = -(-self.My // 16)
        n_yg = -(-tiles // 8)
        waves = n_ang * n_yg * (-(-mb["n_nt"] // self.mfma_variant["fwd"]))
        return int(max(1, min(32, mb["n_kc"], -(-int(self.mfma_variant.get("fwd_waves", 2048)) // waves))))

    def _csr_split(self, n_ang, kb):
        """k split of the CSR forward: csr_variant["ksplit"], or "auto": enough waves for 32 per CU on 256 CUs (a
        thread per (row, 16 detector positions) leaves too few otherwise), at most 16."""
        v = self.csr_variant["ksplit"]
        if v != "auto":
            return int(max(1, min(int(v), max(kb, 1))))
        lanes = -(-self.N_seg // 256) * 256 * -(-self.My // 16) * n_ang
        return int(max(1, min(16, -(-(8192 * 64) // lanes), max(kb // 256, 1))))

    def _mfma_part(self, n):
        p = getattr(self, "_mfma_part_buf", None)
        if p is None or p.numel() < n:
            p = torch.empty(n, dtype=torch.float32, device=self.device)
            self._mfma_part_buf = p
            self.scratch_generation = getattr(self, "scratch_generation", 0) + 1
        return p

    def _dense_forward(self, ib, sino, data, subset, accumulate):
        """Dense PF path: data[r] (+)= sino[r] (My x Kb) @ PF[r] (Kb x CP) for the call's angles."""
        b = self.batches[ib]
        n_ang, om, _, _ = self._angle_set(subset)
        if self.mfma_batches is not None:
            mb = self.mfma_batches[ib]
            ks = self._mfma_split(n_ang, mb)
            part = self._mfma_part(ks * n_ang * self.My * self.N_seg) if ks > 1 else None
            _lib.call("dfm_mfma_forward", _lib.ptr(sino), b["kstride"], b["K_batch"], _lib.ptr(mb["fwd_hi"]),
                      _lib.ptr(mb["fwd_lo"]), _lib.ptr(om), n_ang, self.My, self.N_seg, mb["n_nt"], mb["n_kc"],
                      _lib.ptr(data), _lib.ptr(part), ks, int(accumulate), self.mfma_variant["fwd"], self._stream())
            return
        # FP32: dfm_dense_f32 on the stored (R, kstride, CP) batch, the subset's angles through the angle list
        ks, CP, My = b["kstride"], self.N_seg, self.My
        # the long reduction over the orientations in parts (at least 4096 workgroups, parts of >= 512 terms),
        # summed in order by a second kernel
        tiles = -(-CP // 64) * -(-My // 64) * n_ang
        sp = int(max(1, min(-(-4096 // tiles), b["K_batch"] // getattr(self, "dense_split_min", 512), 65535 // n_ang)))
        part = self._mfma_part(sp * n_ang * My * CP) if sp > 1 else None
        _lib.call("dfm_dense_f32", 0, _lib.ptr(sino), My * ks, ks, 0, _lib.ptr(self.dense_batches[ib]), ks * CP, CP,
                  _lib.ptr(data), My * CP, CP, 1, _lib.ptr(om), n_ang, My, CP, b["K_batch"], 1.0, int(accumulate),
                  sp, _lib.ptr(part), self._stream())

    def _dense_adjoint(self, ib, data, sino, subset, alpha):
        b = self.batches[ib]
        n_ang, om, _, _ = self._angle_set(subset)
        if self.mfma_batches is not None:
            mb = self.mfma_batches[ib]
            _lib.call("dfm_mfma_adjoint", _lib.ptr(data), _lib.ptr(mb["adj_hi"]), _lib.ptr(mb["adj_lo"]),
                      _lib.ptr(om), n_ang, self.My, self.N_seg, mb["n_kt"], mb["n_jc"], _lib.ptr(sino),
                      b["kstride"], float(alpha), self.mfma_variant["adj"], self._stream())
            return
        ks, CP, My = b["kstride"], self.N_seg, self.My
        _lib.call("dfm_dense_f32", 1, _lib.ptr(data), My * CP, CP, 1, _lib.ptr(self.dense_batches[ib]), ks * CP, CP,
                  _lib.ptr(sino), My * ks, ks, 0, _lib.ptr(om), n_ang, My, ks, CP, float(alpha), 0, 1, None,
                  self._stream())

    def direct_cl(self, coeffs, data, subset=None):
        """In-place forward: data = PF @ Radon(coeffs); with subset=m only the rows of subset m's angles are
        written, the others are left as they are (single_phase_forward_operator.py:1646-1729)."""
        check_device(coeffs, self.coeff_shape, "coeffs", self.device)
        check_device(data, self.data_shape, "data", self.device)
        n_ang, _, geo, _ = self._angle_set(subset)
        tab = self._tables(subset)
        sc = self._scratch_buffers(n_ang)
        npix = self.Nx * self.Ny
        st = self._stream()
        for ib, b in enumerate(self.batches):
            ks = b["kstride"]
            _lib.call("dfm_gather_k", _lib.ptr(coeffs), _lib.ptr(sc["img"]), npix, b["k_start"], b["K_batch"], ks, st)
            self._radon_forward(sc["img"], ks, 0, ks, sc["sino"], geo, n_ang, tab)
            self._pf_forward(ib, sc["sino"], data, subset, accumulate=ib > 0)

    def adjoint_cl(self, data, coeffs, subset=None):
        """In-place adjoint: coeffs = A^T data (with subset=m: A_m^T, reading only subset m's rows)
        (single_phase_forward_operator.py:1800-1817, 1966-2033)."""
        check_device(data, self.data_shape, "data", self.device)
        check_device(coeffs, self.coeff_shape, "coeffs", self.device)
        n_ang, _, _, geo = self._angle_set(subset)
        tab = self._tables(subset)
        sc = self._scratch_buffers(n_ang)
        npix = self.Nx * self.Ny
        st = self._stream()
        for ib, b in enumerate(self.batches):
            ks = b["kstride"]
            self._pf_adjoint(ib, data, sc["sino"], subset)
            self._radon_backward(sc["sino"], geo, tab, n_ang, sc["img"], ks, 0, ks)
            _lib.call("dfm_scatter_k", _lib.ptr(sc["img"]), _lib.ptr(coeffs), npix, b["k_start"], b["K_batch"], ks, st)

    # channel-fastest forms (used by the solver): x_k (Npix, K_stride), batch b at columns koff .. koff + kstride
    def direct_k(self, x_k, data, subset=None):
        """data (rows of subset) = A x for a channel-fastest coefficient array."""
        n_ang, _, geo, _ = self._angle_set(subset)
        tab = self._tables(subset)
        sc = self._scratch_buffers(n_ang)
        for ib, b in enumerate(self.batches):
            self._radon_forward(x_k, self.K_stride, b["koff"], b["kstride"], sc["sino"], geo, n_ang, tab)
            self.mark("radon_forward")
            self._pf_forward(ib, sc["sino"], data, subset, accumulate=ib > 0)
            self.mark("pf_forward")

    def mark(self, name):
        """With self.marks a list (a profiling pass), record a HIP event on the stream the kernels run on and append
        (name, event): the launches since the previous mark belong to `name`."""
        if self.marks is not None:
            e = torch.cuda.Event(enable_timing=True)
            e.record(self.queue)
            self.marks.append((name, e))

    def adjoint_k(self, data, out_k, subset=None):
        """out_k = A^T data (channel-fastest)."""
        n_ang, _, _, geo = self._angle_set(subset)
        tab = self._tables(subset)
        sc = self._scratch_buffers(n_ang)
        for ib, b in enumerate(self.batches):
            self._pf_adjoint(ib, data, sc["sino"], subset)
            self._radon_backward(sc["sino"], geo, tab, n_ang, out_k, self.K_stride, b["koff"], b["kstride"])

    def to_k(self, coeffs, out_k=None):
        """(K, Ny, Nx) -> channel-fastest (Npix, K_stride)."""
        npix = self.Nx * self.Ny
        if out_k is None:
            out_k = torch.empty(npix * self.K_stride, dtype=torch.float32, device=self.device)
        for b in self.batches:
            # gather into the batch's columns: a (Npix, kstride) view is not contiguous, so gather per batch into
            # a scratch and copy when there are several batches
            if len(self.batches) == 1:
                _lib.call("dfm_gather_k", _lib.ptr(coeffs), _lib.ptr(out_k), npix, b["k_start"], b["K_batch"],
                          b["kstride"], self._stream())
            else:
                sc = self._scratch_buffers(1)
                _lib.call("dfm_gather_k", _lib.ptr(coeffs), _lib.ptr(sc["img"]), npix, b["k_start"], b["K_batch"],
                          b["kstride"], self._stream())
                out_k.view(npix, self.K_stride)[:, b["koff"]:b["koff"] + b["kstride"]].copy_(
                    sc["img"][: npix * b["kstride"]].view(npix, b["kstride"]))
        return out_k

    def from_k(self, x_k, coeffs):
        """channel-fastest (Npix, K_stride) -> (K, Ny, Nx)."""
        npix = self.Nx * self.Ny
        for b in self.batches:
            if len(self.batches) == 1:
                _lib.call("dfm_scatter_k", _lib.ptr(x_k), _lib.ptr(coeffs), npix, b["k_start"], b["K_batch"],
                          b["kstride"], self._stream())
            else:
                sc = self._scratch_buffers(1)
                sc["img"][: npix * b["kstride"]].view(npix, b["kstride"]).copy_(
                    x_k.view(npix, self.K_stride)[:, b["koff"]:b["koff"] + b["kstride"]])
                _lib.call("dfm_scatter_k", _lib.ptr(sc["img"]), _lib.ptr(coeffs), npix, b["k_start"], b["K_batch"],
                          b["kstride"], self._stream())
        return coeffs

    # ================================================================== allocating forms
    def direct(self, coeffs):
        """The forward operator, allocating its output: a torch tensor in gives a torch tensor; a NumPy array gives
        a NumPy array (single_phase_forward_operator.py:1107-1127)."""
        if isinstance(coeffs, torch.Tensor):
            data = torch.empty(self.data_shape, dtype=torch.float32, device=self.device)
            self.direct_cl(coeffs, data)
            return data
        x = torch.from_numpy(as_host(coeffs, self.coeff_shape, "coeffs")).to(self.device)
        data = torch.empty(self.data_shape, dtype=torch.float32, device=self.device)
        self.direct_cl(x, data)
        return data.cpu().numpy()

    def adjoint(self, data):
        """The adjoint, allocating its output (single_phase_forward_operator.py:1777-1796)."""
        if isinstance(data, torch.Tensor):
            coeffs = torch.empty(self.coeff_shape, dtype=torch.float32, device=self.device)
            self.adjoint_cl(data, coeffs)
            return coeffs
        d = torch.from_numpy(as_host(data, self.data_shape, "data")).to(self.device)
        out = torch.empty(self.coeff_shape, dtype=torch.float32, device=self.device)
        self.adjoint_cl(d, out)
        return out.cpu().numpy()


def _dot64(a, b):
    """float64 dot product of two float32 device tensors (dfm_dot_partial)."""
    part = torch.empty(2 * 2048, dtype=torch.float64, device=a.device)
    n = _lib.ctypes.c_int32(0)
    _lib.call("dfm_dot_partial", _lib.ptr(a), _lib.ptr(b), a.numel(), _lib.ptr(part), _lib.ctypes.byref(n),
              _lib.stream_handle())
    return float(part[: n.value].sum().item())


def estimate_L_power(op, niter: int = 20, seed: int = 0, eps: float = 1e-30, verbose: int = 1) -> float:
    """Lipschitz constant L = ||A^T A|| by power iteration with float64 dot products
    (single_phase_forward_operator.py:2136-2206): the start is drawn as (Nx, Ny, K) standard normals and transposed
    to (K, Ny, Nx)."""
    if niter < 1:
        raise ValueError("niter must be >= 1")
    rng = np.random.default_rng(seed)
    xh = np.ascontiguousarray(rng.standard_normal((op.Nx, op.Ny, op.K)).astype(np.float32).transpose(2, 1, 0))
    x = torch.from_numpy(xh).to(op.device)
    Ax = torch.empty(op.data_shape, dtype=torch.float32, device=op.device)
    z = torch.empty(op.coeff_shape, dtype=torch.float32, device=op.device)
    xn = np.sqrt(_dot64(x, x)) + eps
    _lib.call("dfm_scale", _lib.ptr(x), float(1.0 / xn), x.numel(), _lib.stream_handle())
    L = 0.0
    for it in range(niter):
        op.direct_cl(x, Ax)
        op.adjoint_cl(Ax, z)
        L = _dot64(x, z) / (_dot64(x, x) + eps)
        zn = np.sqrt(_dot64(z, z)) + eps
        _lib.call("dfm_scale", _lib.ptr(z), float(1.0 / zn), z.numel(), _lib.stream_handle())
        x, z = z, x
        if verbose:
            print(f"[power {it + 1:02d}] L_est={L:.6e}  ||z||={zn:.6e}")
    return float(L)
