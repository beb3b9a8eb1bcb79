"""The single-phase operator with a local basis (operators/local_forward_operator.py:1-532): orientation k carries
coefficients only where a LocalSupport allows it, stored as a flat float32 vector of rectangular patches
(local_support.build_patches), instead of a (K, Ny, Nx) image stack.

MI355X design. The detector, the PF matrix (stored block-sparse / CSR, or generated per call) and its products are
those of SinglePhaseForwardOperator, on the used orientations only, numbered along a Z-order curve through their
voxels so that an orientation batch covers one part of the slice. Only the projector changes: a patch projects onto
a strip of detector bins a few wider than the patch at every angle (csrc/local_basis.hip), so the projection work and
the coefficient memory scale with the patches' area, not with K x Ny x Nx:

    forward  = PF product of the batch sinogram assembled from the strips (zero outside them)
    adjoint  = backprojection of every allowed coefficient from the strip bins of the PF adjoint's sinogram

The strip arithmetic is the native projector's (the same footprint, geometry and scales), so the operator equals
SinglePhaseForwardOperator applied to the coefficients scattered into (K, Ny, Nx) (zero outside the support), up to
float32 summation order. SPDHGPoisson and FISTAPoisson run on it (every array on the GPU; SPDHG fuses its primal
update into the patch backprojection); the support constraint is part of the operator.
"""
from __future__ import annotations

import time

import numpy as np
import torch

from . import _lib
from .grid import MatrixGrid
from .local_support import LocalSupport, build_patches, morton_code, patch_neighbours
from .operator import SinglePhaseForwardOperator, as_host, check_device, event_timer
from .support import fov_support_mask


class LocalSinglePhaseForwardOperator(SinglePhaseForwardOperator):
    """data = PF @ Radon(coeffs) with a local basis (local_forward_operator.py:85-199).

    Coefficients: a flat float32 vector of n_coef (coeff_shape = (n_coef,)); to_dense / from_dense convert from and
    to (K, Ny, Nx) arrays in the grid's orientation order, to_coo gives (k, iy, ix, value) per allowed pair.

    support: a LocalSupport for the grid's K orientations on the (Ny, Nx) grid, cut to the field of view unless
    restrict_to_fov=False; orientations without a voxel are left out. split_patches, merge_gap, min_box_fill: how
    the pairs are grouped into patches (build_patches). spatial_order: number the orientations along a Z-order curve
    through their voxels' centroids (default) instead of the grid's order (internal: the coefficients and to_dense /
    to_coo keep the grid's order). pf_chunk is accepted for the reference's signature; the PF products here tile
    themselves. Only the native projector and the sparse / generated PF modes apply."""

    coefficient_layout = "local"

    def __init__(self, cfg, material, grid, max_gb, support, verbose=False, split_patches=True, merge_gap=2,
                 min_box_fill=0.5, restrict_to_fov=True, spatial_order=True, pf_chunk=None, **kwargs):
        t0 = time.perf_counter()
        projector = kwargs.pop("projector", "native")
        if projector not in ("native", "auto"):
            raise ValueError("LocalSinglePhaseForwardOperator has its own projector (the native discretisation); "
                             f"projector={projector!r} is not supported")
        if kwargs.get("pf_mode", "auto") == "dense":
            raise ValueError("LocalSinglePhaseForwardOperator needs a sparse PF mode ('auto', 'sparse', 'generated')")
        kwargs.setdefault("sparse_max_fill", 1.0)
        kwargs.setdefault("sparse_max_fill_partial", 1.0)
        c = {**cfg, **kwargs}
        Nx, Ny, My = int(c["Nx"]), int(c["Ny"]), int(c["My"])
        mats, sigmas = grid.leaf_arrays()
        mats = np.asarray(mats, np.float64)
        sigmas = np.broadcast_to(np.asarray(sigmas, np.float64), (len(mats),))
        if not isinstance(support, LocalSupport):
            raise TypeError("support must be a LocalSupport")
        if support.K != len(mats) or (support.Ny, support.Nx) != (Ny, Nx):
            raise ValueError(f"support is for K = {support.K} on {support.Ny} x {support.Nx}; the grid has "
                             f"{len(mats)} orientations, the geometry {Ny} x {Nx} voxels")
        if restrict_to_fov:
            R = int(c["N_Omega"])
            ar = np.array(c["angle_range"], np.float64) / 180 * np.pi
            angles = np.linspace(ar[0], ar[1], R, endpoint=False) + (ar[1] - ar[0]) / R / 2
            vs = c.get("voxel_size")
            fov = fov_support_mask(Nx, Ny, My, angles=angles,
                                   image_width=Nx if vs is None else float(vs) * max(Nx, Ny), detector_width=My,
                                   detector_shift=float(c["cor_offset"]))
            support = support.restrict(fov)
        if support.n_pairs == 0:
            raise ValueError("the support is empty (inside the field of view)")
        self.support = support
        self.K_basis = len(mats)
        used = support.orientations_used()
        k_used = np.searchsorted(used, support.k)
        if spatial_order:
            cnt = np.bincount(k_used, minlength=used.size)
            cy = np.bincount(k_used, weights=support.iy, minlength=used.size) / cnt
            cx = np.bincount(k_used, weights=support.ix, minlength=used.size) / cnt
            order = np.argsort(morton_code(cy.astype(np.int64), cx.astype(np.int64)), kind="stable")
        else:
            order = np.arange(used.size)
        rank = np.empty(used.size, np.int64)
        rank[order] = np.arange(used.size)
        self.orientation_index = used[order]                 # operator orientation -> grid orientation
        k_op = rank[k_used]
        perm = np.lexsort((support.ix, support.iy, k_op))    # the pairs sorted by operator orientation
        P = build_patches(k_op[perm], support.iy[perm], support.ix[perm], used.size, split=split_patches,
                          merge_gap=merge_gap, min_box_fill=min_box_fill)
        pair_index = np.empty_like(P["pair_index"])
        pair_index[perm] = P["pair_index"]                  # in the order of the support's pairs
        P["pair_index"] = pair_index
        self._patches = P
        self.n_coef = int(P["n_coef"])
        self.n_patches = int(P["pk"].size)
        self.pf_chunk = pf_chunk
        t_patch = time.perf_counter() - t0
        inner = MatrixGrid(mats[self.orientation_index], sigmas[self.orientation_index])
        super().__init__(cfg, material, inner, max_gb, verbose=False, projector="native", **kwargs)
        if self.pf_mode == "dense":
            self.free_memory()
            raise ValueError("LocalSinglePhaseForwardOperator: the PF matrix could not be set up sparse")
        if self.n_coef >= 2 ** 31:
            raise ValueError(f"{self.n_coef} coefficients: more than 2^31 - 1")
        # stage patches (strip projection) and strips (backprojection) in LDS where they fit; False: the global-memory
        # forms throughout (the same results)
        self.local_staging = True
        self._setup_local()
        self.setup_s = time.perf_counter() - t0
        if verbose:
            print(f"Local basis: {support.n_pairs} pairs ({100 * support.fill:.2f} % of K x Ny x Nx), "
                  f"{self.K} of {self.K_basis} orientations, {self.n_patches} patches, {self.n_coef} coefficients, "
                  f"strips {100 * self.strip_fraction:.1f} % of the sinograms; patches {t_patch:.1f} s, "
                  f"set up {self.setup_s:.1f} s")

    # ---------------------------------------------------------------- base-class hooks
    def _radon_tables(self, idx):
        """No geometry tables: the strip kernels evaluate the footprint from the geometry records."""
        return None

    def _scratch_buffers(self, n_angles):
        """The batch sinogram (n_angles, My, K_batch_max) only (no image batch)."""
        need = max(n_angles, self.N_Omega) * self.My * self.K_batch_max
        sc = getattr(self, "_scratch", None)
        if sc is None or sc["sino"].numel() < need:
            sc = {"sino": torch.empty(need, dtype=torch.float32, device=self.device), "img": None}
            self._scratch = sc
            self.scratch_generation = getattr(self, "scratch_generation", 0) + 1
        return sc

    @property
    def coeff_shape(self):
        return (self.n_coef,)

    # ---------------------------------------------------------------- set-up
    def _setup_local(self):
        """Per orientation batch: the patches' boxes and offsets, the strip of every (angle, patch) (first bin and the
        corner's detector coordinate relative to it, float64 here), the strip bins' patches, the patches of every
        orientation and the coefficients' patches, as device arrays (local_forward_operator.py:264-343)."""
        Pt = self._patches
        R = self.N_Omega
        g = self._geo64
        X, Y, T0, hw = (np.asarray(g[k], np.float64) for k in ("X", "Y", "T0", "hw"))
        ntap = int(self.n_taps)
        dev = self.device
        i32 = lambda a: torch.from_numpy(np.ascontiguousarray(a, dtype=np.int32)).to(dev)
        bounds = np.searchsorted(Pt["pk"], [b["k_start"] for b in self.batches] + [self.K])
        self.local_batches = []
        max_strip = 1
        strip_bins = 0
        for ib, b in enumerate(self.batches):
            p0, p1 = int(bounds[ib]), int(bounds[ib + 1])
            if p1 == p0:
                self.local_batches.append(None)
                continue
            y0, x0, h, w = (Pt[n][p0:p1].astype(np.int64) for n in ("y0", "x0", "h", "w"))
            nb = p1 - p0
            # the detector coordinate of the box corners at every angle, (R, nb)
            ax0, ax1 = np.outer(X, x0), np.outer(X, x0 + w - 1)
            ay0, ay1 = np.outer(Y, y0), np.outer(Y, y0 + h - 1)
            tmin = np.minimum(ax0, ax1) + np.minimum(ay0, ay1) + T0[:, None]
            tmax = np.maximum(ax0, ax1) + np.maximum(ay0, ay1) + T0[:, None]
            # bins a footprint can reach: floor(t - hw) + 1 .. floor(t - hw) + ntap; one bin of margin each side
            off = np.floor(tmin - hw[:, None]) - 1
            last = np.floor(tmax - hw[:, None]) + ntap + 1
            Wp = (last - off + 1).max(axis=0).astype(np.int64)
            pw_off = np.concatenate([[0], np.cumsum(Wp)[:-1]])
            SW = int(Wp.sum())
            if R * SW >= 2 ** 31:
                raise ValueError("too many strip bins in one orientation batch for 32-bit indices; lower max_gb")
            tb = X[:, None] * x0[None] + Y[:, None] * y0[None] + T0[:, None] - off
            ptab = np.empty((R, nb, 2), np.int32)
            ptab[..., 0] = off.astype(np.int32)
            ptab[..., 1] = tb.astype(np.float32).view(np.int32)
            pk = Pt["pk"][p0:p1] - b["k_start"]
            size = h * w
            coff = Pt["coff"][p0:p1].astype(np.int64)
            nblk = -(-size // 256)
            blk = np.stack([np.repeat(np.arange(nb), nblk),
                            np.concatenate([np.arange(c) * 256 for c in nblk])], 1)
            self.local_batches.append(dict(
                n=nb, SW=SW, max_area=int(size.max()), w_max=int(Wp.max()), blk=i32(blk), n_blk=int(nblk.sum()),
                box=i32(np.stack([y0, x0, h, w], 1)),
                coff=torch.from_numpy(coff).to(dev), W=i32(Wp), pw_off=i32(pw_off), ptab=i32(ptab),
                jp=i32(np.repeat(np.arange(nb), Wp)), pk=i32(pk),
                kp_ptr=i32(np.searchsorted(pk, np.arange(b["K_batch"] + 1))),
                pid=i32(np.repeat(np.arange(nb), size)), c_begin=int(coff[0]), n_c=int(size.sum())))
            max_strip = max(max_strip, R * SW)
            strip_bins += SW * R
        self.strip_fraction = strip_bins / float(R * self.My * self.K)
        # the most pixels one bin's footprint covers along a patch row, ceil(2 hw / A) over the angles (float32 as the
        # kernels compute it); 0 (no bound) beyond 4
        A = np.maximum(np.abs(X.astype(np.float32)), np.abs(Y.astype(np.float32))).astype(np.float64)
        nr = int(np.ceil(2.0 * float(np.max(hw / A)) - 1e-9))
        self._n_reach = nr if nr <= 4 else 0
        self._strips = torch.empty(max_strip, dtype=torch.float32, device=dev)
        self._mask = torch.from_numpy(Pt["mask"]).to(dev)
        self._all_omegas = torch.arange(R, dtype=torch.int32, device=dev)

    # ---------------------------------------------------------------- products
    def _local_angles(self, subset):
        if subset is None:
            return self.N_Omega, self._all_omegas
        m = self._subset(subset)
        return len(self.omega_subsets[m]), self._sub_omegas[m]

    def _local_forward(self, x, data, subset):
        n_ang, om = self._local_angles(subset)
        sc = self._scratch_buffers(n_ang)
        st = self._stream()
        first = True
        for ib, lb in enumerate(self.local_batches):
            if lb is None:
                continue
            b = self.batches[ib]
            _lib.call("dfm_local_strips", _lib.ptr(x), _lib.ptr(self._mask), _lib.ptr(lb["box"]), _lib.ptr(lb["coff"]),
                      _lib.ptr(lb["W"]), _lib.ptr(lb["pw_off"]), lb["SW"], _lib.ptr(lb["jp"]), _lib.ptr(lb["ptab"]),
                      lb["n"], lb["max_area"] if self.local_staging else 1 << 30, self._n_reach, _lib.ptr(self.geo_fwd), _lib.ptr(om), n_ang, float(self.radon_scale),
                      _lib.ptr(self._strips), st)
            _lib.call("dfm_local_assemble", _lib.ptr(self._strips), lb["SW"], _lib.ptr(lb["pw_off"]), _lib.ptr(lb["W"]),
                      _lib.ptr(lb["ptab"]), lb["n"], _lib.ptr(lb["kp_ptr"]), b["K_batch"], _lib.ptr(om), n_ang, self.My,
                      b["kstride"], _lib.ptr(sc["sino"]), st)
            self.mark("radon_forward")
            self._pf_forward(ib, sc["sino"], data, subset, accumulate=not first)
            self.mark("pf_forward")
            first = False
        if first:
            if subset is None:
                data.zero_()
            else:
                self.zero_subset_rows(data, subset)

    def _local_adjoint(self, data, subset, mode, out=None, x=None, z=None, T=None, inv_p=0.0, lam=0.0):
        """mode 0: out = A^T data (A_m^T with subset m); mode 1: the SPDHG primal update with dz = A_m^T data."""
        n_ang, om = self._local_angles(subset)
        sc = self._scratch_buffers(n_ang)
        st = self._stream()
        for ib, lb in enumerate(self.local_batches):
            if lb is None:
                continue
            b = self.batches[ib]
            self._pf_adjoint(ib, data, sc["sino"], subset)
            self.mark("pf_adjoint")
            _lib.call("dfm_local_extract", _lib.ptr(sc["sino"]), n_ang, self.My, b["kstride"], _lib.ptr(om),
                      _lib.ptr(lb["ptab"]), lb["n"], _lib.ptr(lb["pk"]), _lib.ptr(lb["jp"]), _lib.ptr(lb["pw_off"]),
                      lb["SW"], _lib.ptr(self._strips), st)
            _lib.call("dfm_local_backproject", _lib.ptr(self._strips), lb["SW"], _lib.ptr(lb["pw_off"]),
                      _lib.ptr(lb["W"]), _lib.ptr(lb["ptab"]), lb["n"], lb["w_max"], _lib.ptr(self.geo_bwd),
                      _lib.ptr(om), n_ang, _lib.ptr(lb["box"]), _lib.ptr(lb["coff"]), _lib.ptr(self._mask),
                      _lib.ptr(lb["pid"]), lb["c_begin"], lb["n_c"], _lib.ptr(lb["blk"]),
                      lb["n_blk"] if self.local_staging else 0,
                      int(self.n_taps), int(mode), _lib.ptr(out), _lib.ptr(x), _lib.ptr(z), _lib.ptr(T),
                      float(inv_p), float(lam), st)
            self.mark("radon_backward_spdhg" if mode == 1 else "radon_backward")

    def direct_cl(self, coeffs, data, subset=None):
        """data = PF @ Radon(coeffs) in place; coeffs (n_coef,) float32; with subset=m only the rows of subset m's
        angles are written (local_forward_operator.py:394-410)."""
        check_device(coeffs, self.coeff_shape, "coeffs", self.device)
        check_device(data, self.data_shape, "data", self.device)
        self._local_forward(coeffs, data, subset)

    def adjoint_cl(self, data, coeffs, subset=None):
        """coeffs = A^T data in place, zero outside the support; subset=m: A_m^T (local_forward_operator.py:412-425)."""
        check_device(data, self.data_shape, "data", self.device)
        check_device(coeffs, self.coeff_shape, "coeffs", self.device)
        self._local_adjoint(data, subset, 0, out=coeffs)

    # the solvers' resident layout is the flat vector itself
    def direct_k(self, x_k, data, subset=None):
        self._local_forward(x_k, data, subset)

    def adjoint_k(self, data, out_k, subset=None):
        self._local_adjoint(data, subset, 0, out=out_k)

    def adjoint_update(self, data, subset, x, z, T, inv_p, lam):
        """The SPDHG primal update with dz = A_m^T data fused into the patch backprojection (x, z, T flat)."""
        self._local_adjoint(data, subset, 1, x=x, z=z, T=T, inv_p=inv_p, lam=lam)

    def to_k(self, coeffs, out_k=None):
        if out_k is None:
            return coeffs.reshape(-1).clone()
        out_k.copy_(coeffs.reshape(-1))
        return out_k

    def from_k(self, x_k, coeffs):
        coeffs.reshape(-1).copy_(x_k)
        return coeffs

    def direct(self, coeffs):
        """The forward operator, allocating its output: NumPy (n_coef,) -> NumPy data, torch -> torch."""
        if isinstance(coeffs, torch.Tensor):
            data = torch.empty(self.data_shape, dtype=torch.float32, device=self.device)
            self.direct_cl(coeffs, data)
            return data
        x = torch.from_numpy(as_host(coeffs, self.coeff_shape, "coeffs")).to(self.device)
        data = torch.empty(self.data_shape, dtype=torch.float32, device=self.device)
        self.direct_cl(x, data)
        return data.cpu().numpy()

    def adjoint(self, data):
        """The adjoint, allocating its output: NumPy data -> NumPy (n_coef,), torch -> torch."""
        if isinstance(data, torch.Tensor):
            out = torch.empty(self.coeff_shape, dtype=torch.float32, device=self.device)
            self.adjoint_cl(data, out)
            return out
        d = torch.from_numpy(as_host(data, self.data_shape, "data")).to(self.device)
        out = torch.empty(self.coeff_shape, dtype=torch.float32, device=self.device)
        self.adjoint_cl(d, out)
        return out.cpu().numpy()

    def autotune(self, subset=0, reps=5):
        """The PF products' variants (the strip kernels have a single form)."""
        if self.omega_subsets is None:
            self.set_omega_subsets(10)
        n_ang, _ = self._local_angles(subset)
        sc = self._scratch_buffers(n_ang)
        g = torch.Generator(device=self.device)
        g.manual_seed(0)
        out = {}
        self._autotune_pf(subset, sc, g, event_timer(self.device, reps), out)
        self.tuning = out
        return out

    def tune_projector(self, *args, **kwargs):
        return {}

    def rebuild(self, pf=False, geometry=False, projector=False):
        raise NotImplementedError("the local-basis operator does not support parameter updates (rebuild)")

    # ---------------------------------------------------------------- coefficients
    def _pairs(self):
        s = self.support
        return s.k, s.iy, s.ix, self._patches["pair_index"]

    def patch_neighbours(self):
        """(right, down, left, up) neighbour index arrays of the flat coefficients (local_support.patch_neighbours)."""
        return patch_neighbours(self._patches)

    def to_dense(self, x):
        """(K_basis, Ny, Nx) float32 in the grid's orientation order, zero outside the support."""
        x = x.detach().cpu().numpy() if isinstance(x, torch.Tensor) else np.asarray(x)
        k, iy, ix, pi = self._pairs()
        out = np.zeros((self.K_basis, self.Ny, self.Nx), dtype=np.float32)
        out[k, iy, ix] = x.reshape(-1)[pi]
        return out

    def from_dense(self, coeffs):
        """The flat coefficients of a (K_basis, Ny, Nx) array (values outside the support dropped)."""
        c = np.asarray(coeffs)
        if c.shape != (self.K_basis, self.Ny, self.Nx):
            raise ValueError(f"coeffs must have shape {(self.K_basis, self.Ny, self.Nx)}, not {c.shape}")
        k, iy, ix, pi = self._pairs()
        x = np.zeros(self.n_coef, dtype=np.float32)
        x[pi] = c[k, iy, ix]
        return x

    def to_coo(self, x):
        """(k, iy, ix, value) of every allowed pair, k in the grid's orientation order."""
        x = x.detach().cpu().numpy() if isinstance(x, torch.Tensor) else np.asarray(x)
        k, iy, ix, pi = self._pairs()
        return k.copy(), iy.copy(), ix.copy(), x.reshape(-1)[pi].astype(np.float32)

    def coefficient_mask(self, mask=None):
        """(n_coef,) float32: 1 for the allowed coefficients (only where the (Ny, Nx) mask is true, if given)."""
        m = self._patches["mask"].astype(np.float32)
        if mask is not None:
            mk = np.asarray(mask, dtype=bool)
            k, iy, ix, pi = self._pairs()
            m[pi[~mk[iy, ix]]] = 0.0
        return m

    # ---------------------------------------------------------------- memory
    def device_bytes(self):
        n = super().device_bytes()
        for lb in getattr(self, "local_batches", None) or []:
            if lb is not None:
                n += sum(v.numel() * v.element_size() for v in lb.values() if isinstance(v, torch.Tensor))
        for name in ("_strips", "_mask"):
            t = getattr(self, name, None)
            if t is not None:
                n += t.numel() * t.element_size()
        return n

    def free_memory(self):
        self.local_batches = None
        self._strips = self._mask = None
        super().free_memory()
