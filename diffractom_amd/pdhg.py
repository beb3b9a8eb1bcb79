"""PDHGPoisson on MI355X: Chambolle-Pock with diagonal preconditioning for the Poisson deviance, with the reference's
signature and iteration (optimization/pdhg_poisson.py:120-265).

Iteration (theta = 1), over all rotation angles:
    y     <- prox_{Sigma F*}(y + Sigma A x_bar)          (float64 arithmetic, dfm_dual_update_sub on every angle)
    x_new <- max(0, x - T (A^T y + lam));  x_bar <- 2 x_new - x
    A x   <- (A x_bar + A x) / 2                          (recomputed from x every 50 iterations: float32 drift)
with T = tau0 / (A^T 1_counted) on the support, Sigma = sigma0 / (A 1_S), tau0 sigma0 = 1, rho = sigma0 / tau0
("auto": 3e4 / mean(b)^2 over the counted points).

MI355X design: x, x_bar and T stay on the GPU for the whole run in the orientation-fastest layout (Npix, K) the Radon
kernels read, so an iteration is the dual update, the PF adjoint, one backprojection launch that also applies the
primal step and the extrapolation (mode 3 of dfm_radon_backward_table: no A^T y array), the forward of x_bar and the
average. Backprojection choices without that mode backproject into a scratch image and update with
dfm_pdhg_primal_update. The host-streamed variant of the reference (NumPy x0 with x in host memory) is not needed:
every array fits the HBM; a NumPy x0 is uploaded, updated in place and returned.
"""
from __future__ import annotations

import ctypes
import time

import numpy as np
import torch

from . import _lib
from .operator import as_host, check_device
from .spdhg import DFM_MAX_PARTIALS, RHO_AUTO, _periodic, support_to_device


class PDHGPoisson:
    """Chambolle-Pock for the Poisson deviance (pdhg_poisson.py:120-146)."""

    def __init__(self, operator, prox_kind="nonneg", lam=0.0, support="fov", eps_counts=1e-3, rho="auto"):
        """rho = sigma0 / tau0 (tau0 sigma0 = 1); "auto": RHO_AUTO / mean(b)^2 over the counted data points; a
        number sets it. prox_kind "nonneg" or "nonneg_l1" (lam sum(x) with lam > 0)."""
        if prox_kind not in ("nonneg", "nonneg_l1"):
            raise ValueError("PDHGPoisson supports prox_kind 'nonneg' and 'nonneg_l1'")
        if getattr(operator, "coefficient_layout", "dense") == "local":
            raise NotImplementedError("PDHGPoisson runs on the (K, Ny, Nx) coefficient layout; with a local basis use "
                                      "SPDHGPoisson or FISTAPoisson")
        self.op = operator
        self.lam = float(lam) if prox_kind == "nonneg_l1" else 0.0
        self.eps_counts = float(eps_counts)
        self.rho = rho if rho == "auto" else float(rho)
        self.support_mask = support_to_device(operator, support)
        self.state = None
        self.iter_stats = []
        self.final_stats = None
        self._part = torch.zeros(2 * DFM_MAX_PARTIALS, dtype=torch.float64, device=operator.device)

    def _deviance(self, mu):
        n = ctypes.c_int32(0)
        _lib.call("dfm_deviance_partial", _lib.ptr(mu), _lib.ptr(self.b), _lib.ptr(self.w), self.pw, _lib.ptr(self.g),
                  self.pg, _lib.ptr(self.bg), self.pbg, float(self.eps_counts), mu.numel(), _lib.ptr(self._part),
                  ctypes.byref(n), self.op._stream())
        return float(self._part[: n.value].sum().item())

    def _mean_counted(self, a):
        n = ctypes.c_int32(0)
        _lib.call("dfm_mean_counted_partial", _lib.ptr(a), _lib.ptr(self.w), self.pw, a.numel(), _lib.ptr(self._part),
                  ctypes.byref(n), self.op._stream())
        v = self._part[: 2 * n.value].view(2, n.value).sum(dim=1).cpu().numpy()
        return float(v[0] / max(v[1], 1.0))

    def _zero_padding(self, a_k):
        """Padding channels (K_batch < kstride) of a channel-fastest array to 0."""
        op = self.op
        v = a_k.view(op.Nx * op.Ny, op.K_stride)
        for bt in op.batches:
            if bt["kstride"] > bt["K_batch"]:
                v[:, bt["koff"] + bt["K_batch"]: bt["koff"] + bt["kstride"]] = 0.0

    def start(self, x0, b, weights=None, gain=None, background=None, state=None):
        """Steps, the resident channel-fastest x, x_bar, T, and A x, A x_bar (pdhg_poisson.py:147-217)."""
        op = self.op
        dev = op.device
        st = op._stream()
        cshape, dshape = tuple(op.coeff_shape), tuple(op.data_shape)
        self.host_x0 = None
        if isinstance(x0, torch.Tensor):
            self.x = check_device(x0, cshape, "x0", dev)
        else:
            self.host_x0 = as_host(x0, cshape, "x0")
            self.x = torch.from_numpy(self.host_x0.copy()).to(dev)
        bh = b if isinstance(b, torch.Tensor) else torch.from_numpy(as_host(b, dshape, "b"))
        if tuple(bh.shape) != dshape:
            raise ValueError(f"b must have shape {dshape}, not {tuple(bh.shape)}")
        self.b = bh.to(dev, torch.float32).contiguous()
        self.w, self.pw = _periodic(weights, dshape, "weights", dev)
        self.g, self.pg = _periodic(gain, dshape, "gain", dev)
        self.bg, self.pbg = _periodic(background, dshape, "background", dev)
        nd = int(np.prod(dshape))
        npix = op.Nx * op.Ny
        nk = npix * op.K_stride
        rho = self.rho
        if rho == "auto":
            mean_b = self._mean_counted(self.b)
            rho = RHO_AUTO / mean_b ** 2 if mean_b > 0 else 1.0
        self.rho_used = float(rho)
        tau0, sigma0 = 1.0 / np.sqrt(rho), np.sqrt(rho)
        supp = self.support_mask
        self.y = torch.zeros(dshape, dtype=torch.float32, device=dev)
        self.sig = torch.empty(dshape, dtype=torch.float32, device=dev)
        self.Axb = torch.empty(dshape, dtype=torch.float32, device=dev)
        self.Ax = torch.empty(dshape, dtype=torch.float32, device=dev)
        self.xk = op.to_k(self.x)
        self.xbk = torch.empty(nk, dtype=torch.float32, device=dev)
        self.Tk = torch.empty(nk, dtype=torch.float32, device=dev)
        self.gradk = None
        # Sigma from the row sums A 1_S, T from the column sums A^T 1_counted (pdhg_poisson.py:199-205)
        ones = self.xbk
        _lib.call("dfm_support_ones_k", _lib.ptr(ones), _lib.ptr(supp), npix, op.K_stride, op.K_stride, st)
        self._zero_padding(ones)
        op.direct_k(ones, self.Axb)
        _lib.call("dfm_dual_steps", _lib.ptr(self.sig), _lib.ptr(self.Axb), _lib.ptr(self.w), self.pw, float(sigma0),
                  nd, st)
        _lib.call("dfm_set_mask_ones", _lib.ptr(self.Axb), _lib.ptr(self.w), self.pw, nd, st)
        colsum = torch.empty(nk, dtype=torch.float32, device=dev)
        op.adjoint_k(self.Axb, colsum)
        _lib.call("dfm_tinit", _lib.ptr(self.Tk), npix, op.K_stride, st)
        _lib.call("dfm_tmin", _lib.ptr(self.Tk), _lib.ptr(colsum), _lib.ptr(supp), npix, op.K_stride, op.K_stride,
                  float(np.float32(tau0)), st)
        _lib.call("dfm_tfinal", _lib.ptr(self.Tk), npix, op.K_stride, op.K_stride, st)
        self._zero_padding(self.Tk)
        del colsum
        _lib.call("dfm_project_support_k", _lib.ptr(self.xk), _lib.ptr(supp), npix, op.K_stride, op.K_stride, st)
        self._zero_padding(self.xk)
        if state is not None:
            op.to_k(torch.from_numpy(as_host(state["x_bar"], cshape, "state['x_bar']")).to(dev), self.xbk)
            self.y.copy_(torch.from_numpy(as_host(state["y"], dshape, "state['y']")).to(dev))
        else:
            self.xbk.copy_(self.xk)
        op.direct_k(self.xk, self.Ax)
        op.direct_k(self.xbk, self.Axb)
        self._all = torch.arange(op.N_Omega, dtype=torch.int32, device=dev)
        self._it0 = int(state["it"]) if state is not None else 0

    def iterate(self):
        """One Chambolle-Pock iteration (pdhg_poisson.py:222-228); A x of the new iterate in self.Ax."""
        op = self.op
        st = op._stream()
        row = op.My * op.N_seg
        # y <- prox(y + Sigma A x_bar); Axb then holds y' - y (not needed: A x_bar is recomputed below)
        _lib.call("dfm_dual_update_sub", _lib.ptr(self.y), _lib.ptr(self.Axb), _lib.ptr(self.sig), _lib.ptr(self.b),
                  _lib.ptr(self.w), self.pw, _lib.ptr(self.g), self.pg, _lib.ptr(self.bg), self.pbg,
                  float(self.eps_counts), _lib.ptr(self._all), op.N_Omega, row, st)
        if op.em_fused(None):
            n_ang, _, _, geo = op._angle_set(None)
            tab = op._tables(None)
            sc = op._scratch_buffers(n_ang)
            for ib, bt in enumerate(op.batches):
                op._pf_adjoint(ib, self.y, sc["sino"], None)
                op._radon_backward(sc["sino"], geo, tab, n_ang, None, op.K_stride, bt["koff"], bt["kstride"],
                                   x=self.xk, z=self.xbk, T=self.Tk, lam=self.lam, pdhg=True)
        else:
            if self.gradk is None:
                self.gradk = torch.empty_like(self.xk)
            op.adjoint_k(self.y, self.gradk)
            _lib.call("dfm_pdhg_primal_update", _lib.ptr(self.xk), _lib.ptr(self.xbk), _lib.ptr(self.gradk),
                      _lib.ptr(self.Tk), float(self.lam), self.xk.numel(), st)
        op.direct_k(self.xbk, self.Axb)
        _lib.call("dfm_average", _lib.ptr(self.Ax), _lib.ptr(self.Axb), self.Ax.numel(), st)

    def coefficients(self):
        """The current iterate in the (K, Ny, Nx) layout."""
        self.op.from_k(self.xk, self.x)
        return self.x

    def run(self, x0, b, niter, weights=None, gain=None, background=None, verbose=0, diagnostics_interval=10,
            callback=None, state=None):
        """niter iterations (pdhg_poisson.py:147-265). x0: a torch tensor on the operator's device or a NumPy array,
        either updated in place and returned; shape op.coeff_shape. b, weights, gain, background: as
        SPDHGPoisson.run. state: the solver.state of an earlier run ({"it", "x_bar", "y"}) to continue it.
        iter_stats: per iteration {"iter", "f", "rho", "time"} (f the deviance of the iterate); final_stats:
        {"niter", "final_f", "setup_s", "rho", "time"}; callback(it, x, stats) after every iteration, a true return
        value stops the run."""
        t_start = time.perf_counter()
        self.start(x0, b, weights, gain, background, state)
        op = self.op
        t_setup = time.perf_counter() - t_start
        self.iter_stats = []
        f = self._deviance(self.Ax)
        for k in range(int(niter)):
            it = self._it0 + k + 1
            self.iterate()
            if it % 50 == 0:                       # float32 drift of the A x recursion
                op.direct_k(self.xk, self.Ax)
            f = self._deviance(self.Ax)
            stats = {"iter": it, "f": f, "rho": self.rho_used, "time": time.perf_counter() - t_start}
            self.iter_stats.append(stats)
            stop = False
            if callback is not None:
                self.coefficients()
                stop = bool(callback(it, self.x if self.host_x0 is None else self.x.cpu().numpy(), stats))
            if verbose and (k == 0 or (k + 1) % diagnostics_interval == 0 or k + 1 == niter or stop):
                print(f"[pdhg {it:5d}] f={f:.8e} {time.perf_counter() - t_start:.1f} s", flush=True)
            if stop:
                break
        xb = torch.empty_like(self.x)
        op.from_k(self.xbk, xb)
        self.state = {"it": self._it0 + len(self.iter_stats), "x_bar": xb.cpu().numpy(), "y": self.y.cpu().numpy()}
        self.final_stats = {"niter": len(self.iter_stats), "final_f": f, "setup_s": t_setup, "rho": self.rho_used,
                            "time": time.perf_counter() - t_start}
        self.coefficients()
        if self.host_x0 is not None:
            self.host_x0[...] = self.x.cpu().numpy()
            return self.host_x0
        return self.x
