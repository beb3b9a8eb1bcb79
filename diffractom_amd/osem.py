"""OSEMPoisson on MI355X: ordered-subsets EM (MLEM with subsets=None) for the Poisson deviance, with the reference's
signature and iteration (optimization/osem_poisson.py:40-185).

Loss on the fitted angles (angles in some subset), mu = max(A x, 0) + bg + eps_counts g:
    f(x) = sum_i w_i / g_i [ mu_i - b_i + b_i log(b_i / mu_i) ]
For the next subset m of a random permutation per epoch (numpy default_rng(seed)):
    x <- x A_m^T(w b / (g mu)) / (A_m^T(w / g) + lam)      (where the sensitivity A_m^T(w / g) > 0; 0 off the support)

MI355X design: x and the per-subset sensitivities stay on the GPU for the whole run in the orientation-fastest layout
(Npix, K) the Radon kernels read (as SPDHGPoisson's x, z, T), so an iteration is the subset's forward projection and
PF product, the EM ratio (one element-wise kernel, dfm_poisson_em_ratio), the subset's PF adjoint and one
backprojection kernel that also applies the multiplicative update (mode 2 of dfm_radon_backward_table): no
gather / scatter passes and no back-projection array. Backprojection choices without that mode (the plain tile,
LDS-resident and arithmetic kernels) backproject into a scratch image and update with dfm_em_update. The support is
applied at the start: a zero coefficient stays zero under the multiplicative update. The deviance is a float64
reduction (dfm_deviance_partial). Angles outside every subset are folded into the weights (w = 0), so every kernel
sees one weight array.
"""
from __future__ import annotations

import ctypes
import time

import numpy as np
import torch

from . import _lib
from .operator import as_host, check_device
from .spdhg import DFM_MAX_PARTIALS, _periodic, support_to_device


class OSEMPoisson:
    """Ordered-subsets EM for the Poisson deviance (optimization/osem_poisson.py:40-80)."""

    def __init__(self, operator, subsets=8, lam=0.0, support="fov", eps_counts=1e-3, seed=0,
                 store_sensitivities=True):
        """subsets: the number of interleaved angle subsets, a list of angle-index arrays (angles in no subset are
        not fitted), or None: MLEM over all angles. lam: weight of the l1 penalty lam sum(x). support: "fov", None
        or an (Ny, Nx) array. seed: the order of the subsets within each epoch. store_sensitivities: keep
        A_m^T(w / g) per subset (False: recompute it every iteration)."""
        if getattr(operator, "coefficient_layout", "dense") == "local":
            raise NotImplementedError("OSEMPoisson runs on the (K, Ny, Nx) coefficient layout; with a local basis use "
                                      "SPDHGPoisson or FISTAPoisson")
        self.op = operator
        if subsets is None:
            self.subsets = None
            self.M = 1
        else:
            self.subsets = operator.set_omega_subsets(subsets)
            self.M = len(self.subsets)
        self.lam = float(lam)
        if self.lam < 0:
            raise ValueError("lam must be >= 0")
        self.eps_counts = float(eps_counts)
        self.rng = np.random.default_rng(seed)
        self.store_sensitivities = bool(store_sensitivities)
        self.mask = support_to_device(operator, support)
        self.pmask = int(self.mask.numel()) if self.mask is not None else 0
        self.iter_stats = []
        self.final_f = None
        self.start_floor_used = None
        self._part = torch.zeros(2 * DFM_MAX_PARTIALS, dtype=torch.float64, device=operator.device)

    def _sub(self, m):
        return None if self.subsets is None else m

    def _adj(self, r, out_k, m):
        """out_k = A_m^T r in the orientation-fastest layout."""
        self.op.adjoint_k(r, out_k, subset=self._sub(m))

    def deviance(self):
        """The deviance of the current iterate on the fitted angles (a full forward)."""
        self.op.direct_k(self.xk, self.pred)
        return self._deviance(self.pred)

    def coefficients(self):
        """The current iterate in the (K, Ny, Nx) layout (the x that start returned, updated)."""
        self.op.from_k(self.xk, self.x)
        return self.x

    def _deviance(self, pred):
        n = ctypes.c_int32(0)
        _lib.call("dfm_deviance_partial", _lib.ptr(pred), _lib.ptr(self.b), _lib.ptr(self.w), self.pw,
                  _lib.ptr(self.g), self.pg, _lib.ptr(self.bg), self.pbg, float(self.eps_counts), pred.numel(),
                  _lib.ptr(self._part), ctypes.byref(n), self.op._stream())
        return float(self._part[: n.value].sum().item())

    def start(self, x0, b, weights=None, gain=None, background=None, start_floor=0.01):
        """The data term (fitted angles folded into the weights), the sensitivities and the start of x
        (osem_poisson.py:111-148); returns x (x0 itself for a torch x0). Then iterate / epoch run the updates."""
        op = self.op
        dev = op.device
        st = op._stream()
        cshape, dshape = tuple(op.coeff_shape), tuple(op.data_shape)
        self.on_gpu = isinstance(x0, torch.Tensor)
        if self.on_gpu:
            x = check_device(x0, cshape, "x0", dev)
        else:
            x = torch.from_numpy(as_host(x0, cshape, "x0").copy()).to(dev)
        nd = int(np.prod(dshape))
        bh = b if isinstance(b, torch.Tensor) else torch.from_numpy(as_host(b, dshape, "b"))
        if tuple(bh.shape) != dshape:
            raise ValueError(f"b must have shape {dshape}, not {tuple(bh.shape)}")
        self.b = bh.to(dev, torch.float32).contiguous()
        self.w, self.pw = _periodic(weights, dshape, "weights", dev)
        self.g, self.pg = _periodic(gain, dshape, "gain", dev)
        self.bg, self.pbg = _periodic(background, dshape, "background", dev)
        if self.subsets is not None:
            fit = np.zeros(dshape[0], bool)
            for s in self.subsets:
                fit[np.asarray(s, np.int64)] = True
            if not fit.all():
                # angles in no subset: weight 0 (poisson_term.py:44-47 multiplies the weight by fit)
                wf = torch.ones(nd, dtype=torch.float32, device=dev) if self.w is None else \
                    self.w.repeat(nd // self.pw)
                rows = torch.from_numpy(np.repeat(fit, nd // dshape[0]).astype(np.float32)).to(dev)
                self.w, self.pw = (wf * rows).contiguous(), nd
        self.pred = torch.zeros(dshape, dtype=torch.float32, device=dev)
        self.r = torch.empty(dshape, dtype=torch.float32, device=dev)
        wg = torch.empty(dshape, dtype=torch.float32, device=dev)
        _lib.call("dfm_poisson_weight_over_gain", _lib.ptr(self.w), self.pw, _lib.ptr(self.g), self.pg, nd,
                  _lib.ptr(wg), st)
        self.wg = wg
        nk = op.Nx * op.Ny * op.K_stride
        self.back = None
        n_sens = self.M if self.store_sensitivities else 1
        self.sens = [torch.empty(nk, dtype=torch.float32, device=dev) for _ in range(n_sens)]
        if self.store_sensitivities:
            for m in range(self.M):
                self._adj(wg, self.sens[m], m)

        # start (osem_poisson.py:125-148)
        n = int(np.prod(cshape))
        xh = x.cpu().numpy().ravel().astype(np.float32)
        allowed = np.resize(self.mask.cpu().numpy().astype(bool), n) if self.pmask else np.ones(n, bool)
        xh = np.where(allowed, xh, 0.0).astype(np.float32)
        if start_floor:
            pos = xh[xh > 0]
            if pos.size:
                floor = float(start_floor) * float(np.median(pos))
            else:
                stot = 0.0
                pix = allowed[: op.Nx * op.Ny]                    # the (Ny, Nx) mask; the same for every k
                for m in range(self.M):
                    if not self.store_sensitivities:
                        self._adj(wg, self.sens[0], m)
                    s = self.sens[m if self.store_sensitivities else 0].view(op.Nx * op.Ny, op.K_stride)
                    stot += float(s[torch.from_numpy(pix).to(dev)].sum(dtype=torch.float64).item())
                wgh = wg.cpu().numpy().ravel().astype(np.float64)
                tot = float(wgh @ self.b.cpu().numpy().ravel().astype(np.float64))
                if self.pbg:
                    tot -= float((wgh * np.resize(self.bg.cpu().numpy().astype(np.float64), nd)).sum())
                floor = max(tot, 0.0) / stot if stot > 0 else 1.0
            xh = np.where(allowed, np.maximum(xh, floor), 0.0).astype(np.float32)
            self.start_floor_used = floor
        x.copy_(torch.from_numpy(xh.reshape(cshape)).to(dev))
        self.x = x
        self.xk = op.to_k(x)
        return x

    def iterate(self, m):
        """One update with subset m (osem_poisson.py:155-166): x <- x A_m^T(w b / (g mu)) / (A_m^T(w / g) + lam)."""
        op, st = self.op, self.op._stream()
        sub = self._sub(m)
        op.direct_k(self.xk, self.pred, subset=sub)
        _lib.call("dfm_poisson_em_ratio", _lib.ptr(self.pred), _lib.ptr(self.b), _lib.ptr(self.w), self.pw,
                  _lib.ptr(self.g), self.pg, _lib.ptr(self.bg), self.pbg, float(self.eps_counts), self.pred.numel(),
                  _lib.ptr(self.r), st)
        if self.store_sensitivities:
            s = self.sens[m]
        else:
            self._adj(self.wg, self.sens[0], m)
            s = self.sens[0]
        if op.em_fused(sub):
            n_ang, _, _, geo = op._angle_set(sub)
            tab = op._tables(sub)
            sc = op._scratch_buffers(n_ang)
            for ib, bt in enumerate(op.batches):
                op._pf_adjoint(ib, self.r, sc["sino"], sub)
                op._radon_backward(sc["sino"], geo, tab, n_ang, None, op.K_stride, bt["koff"], bt["kstride"],
                                   x=self.xk, T=s, lam=self.lam, em=True)
            return
        if self.back is None:
            self.back = torch.empty_like(self.xk)
        op.adjoint_k(self.r, self.back, subset=sub)
        _lib.call("dfm_em_update", _lib.ptr(self.xk), _lib.ptr(self.back), _lib.ptr(s), None, 0, float(self.lam),
                  self.xk.numel(), st)

    def epoch(self):
        """All subsets once, in a random order (osem_poisson.py:154)."""
        order = self.rng.permutation(self.M) if self.M > 1 else [0]
        for m in order:
            self.iterate(int(m))

    def run(self, x0, b, n_epochs, weights=None, gain=None, background=None, start_floor=0.01, verbose=0,
            diagnostics_interval=1, callback=None):
        """n_epochs passes over all subsets (osem_poisson.py:95-185). x0: a torch tensor on the operator's device
        (updated in place and returned) or a NumPy array (a new NumPy array is returned); shape op.coeff_shape.
        b, weights, gain, background: as SPDHGPoisson.run.

        start_floor: every allowed coefficient starts at least at start_floor x the median of the positive allowed
        ones of x0 (0 or None: x0 as given, zeroed outside the support); with no positive allowed coefficient x
        starts flat at the value whose prediction matches the fitted data's total (background subtracted).
        Diagnostics: the deviance on the fitted angles (a full forward) every diagnostics_interval epochs and at
        the end, in iter_stats (epoch, f, time) and final_f; callback(epoch, x, stats) after every epoch, a true
        return value stops the run."""
        t0 = time.perf_counter()
        x = self.start(x0, b, weights, gain, background, start_floor)
        self.iter_stats = []
        f = None
        for ep in range(int(n_epochs)):
            self.epoch()
            stats = {"epoch": ep + 1, "time": time.perf_counter() - t0}
            f = None
            if diagnostics_interval and ((ep + 1) % int(diagnostics_interval) == 0 or ep + 1 == int(n_epochs)):
                f = self.deviance()
                stats["f"] = f
                if verbose:
                    print(f"OSEM epoch {ep + 1}: deviance {f:.6g} ({stats['time']:.1f} s)", flush=True)
            self.iter_stats.append(stats)
            if callback is not None:
                self.coefficients()
                if callback(ep + 1, x if self.on_gpu else x.cpu().numpy(), stats):
                    break
        if f is None:
            f = self.deviance()
        self.final_f = f
        self.coefficients()
        return x if self.on_gpu else x.cpu().numpy()
