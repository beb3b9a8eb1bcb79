"""FISTAPoisson on MI355X: FISTA in the separable-quadratic-surrogate (SQS) diagonal metric with backtracking for the
Poisson deviance, with the reference's signature and iteration (optimization/fista_poisson.py:311-617, the
GPU-resident path).

    D = 1 / A^T(c A 1_S),  c = w / g max(b, m) / m^2,  m = max(max(mu, 0) + bg, floor g)   (the metric, at a prediction)
    x_new = max(0, y - theta D (grad f(y) + lam)),  accepted when f(x_new) <= f(y) + <grad, x_new - y>
            + |x_new - y|^2_{(theta D)^-1} / 2 (else: the metric at y, then theta halves)
    y = x_new + beta (x_new - x) (A y likewise), momentum restart when the objective increases; the metric refreshed
    (and the momentum restarted) after the iterations in `refresh`; from x = 0 the first n_start iterations fit the
    quadratic approximation of the deviance at the data (weighted least squares) instead.

The coefficient arrays stay on the GPU in the (K, Ny, Nx) layout; the per-point and per-coefficient pieces are HIP
kernels (dfm_fista_*: value and gradient, curvature, metric, step, step sums, extrapolation; float64 sums), the
products the operator's direct_cl / adjoint_cl. The reference's host-streamed variant is not needed (every array
fits the HBM): a NumPy x0 is uploaded, updated in place and returned.
"""
from __future__ import annotations

import ctypes
import time

import numpy as np
import torch

from . import _lib
from .operator import as_host, check_device
from .spdhg import DFM_MAX_PARTIALS, _periodic, support_to_device

DEFAULT_REFRESH = (5, 10, 20, 40, 80, 160, 320, 640, 1280, 2560, 5120)   # fista_poisson.py:284


class FISTAPoisson:
    """FISTA in the SQS metric for the Poisson deviance (fista_poisson.py:311-370)."""

    def __init__(self, operator, prox_kind="nonneg", lam=0.0, support="fov", eps_counts=1e-3, floor_counts=1.0,
                 refresh=DEFAULT_REFRESH, refresh_on_failure=True, max_backtracks=60, start="wls", n_start=20):
        if prox_kind not in ("nonneg", "nonneg_l1"):
            raise ValueError("FISTAPoisson supports prox_kind 'nonneg' and 'nonneg_l1' (element-wise proximal "
                             "operators in the diagonal metric)")
        if start not in ("wls", "uniform", "x0"):
            raise ValueError("start must be 'wls', 'uniform' or 'x0'")
        self.op = operator
        self.prox_kind = prox_kind
        self.lam = float(lam) if prox_kind == "nonneg_l1" else 0.0
        self.eps_counts = float(eps_counts)
        self.floor_counts = float(floor_counts)
        self.refresh = tuple(int(r) for r in refresh)
        self.max_backtracks = int(max_backtracks)
        self.refresh_on_failure = bool(refresh_on_failure)
        self.start = start
        self.n_start = int(n_start)
        self._mode = 0
        self.adjoint_scale = 1.0
        self.support_gpu = support_to_device(operator, support)
        self.npix = int(self.support_gpu.numel()) if self.support_gpu is not None else 0
        self.state = None
        self.iter_stats = []
        self.final_stats = None
        self._part = torch.zeros(3 * DFM_MAX_PARTIALS, dtype=torch.float64, device=operator.device)

    # ---- reductions and element-wise pieces
    def _sums(self, n_sums, n):
        return self._part[: n_sums * n].view(n_sums, n).sum(dim=1).cpu().numpy()

    def _data_term(self, mu, write_grad):
        """f at mu (the deviance, or its quadratic approximation at the data in the start phase); write_grad: mu
        becomes df/dmu (fista_poisson.py:376-384)."""
        n = ctypes.c_int32(0)
        _lib.call("dfm_fista_value_grad_partial", _lib.ptr(mu), _lib.ptr(self.b), _lib.ptr(self.w), self.pw,
                  _lib.ptr(self.g), self.pg, _lib.ptr(self.bg), self.pbg, float(self.eps_counts), int(write_grad),
                  int(self._mode), float(self.floor_counts), mu.numel(), _lib.ptr(self._part), ctypes.byref(n),
                  self.op._stream())
        return float(self._sums(1, n.value)[0])

    def _ew(self, op_code, out, a=None, c=None, D=None, th=0.0, lam=0.0):
        _lib.call("dfm_fista_elementwise", op_code, _lib.ptr(out), _lib.ptr(a), _lib.ptr(c), _lib.ptr(D),
                  _lib.ptr(self.support_gpu), self.npix, float(th), float(lam), out.numel(), self.op._stream())

    def _ones_forward(self, scratch, x_tmp):
        """scratch = A 1_S; x_tmp = 1_S (fista_poisson.py:386-391)."""
        self._ew(1, x_tmp)
        self.op.direct_cl(x_tmp, scratch)

    def _metric(self, mu, scratch, x_tmp, D):
        """D = 1 / A^T(c A 1_S) at the prediction mu, and the operator's adjoint scale <A^T u, 1_S> / <u, A 1_S>
        (fista_poisson.py:393-412)."""
        self._ones_forward(scratch, x_tmp)
        n = ctypes.c_int32(0)
        _lib.call("dfm_fista_curvature_partial", _lib.ptr(scratch), _lib.ptr(mu), _lib.ptr(self.b), _lib.ptr(self.w),
                  self.pw, _lib.ptr(self.g), self.pg, _lib.ptr(self.bg), self.pbg, float(self.floor_counts),
                  int(self._mode), scratch.numel(), _lib.ptr(self._part), ctypes.byref(n), self.op._stream())
        u_r = float(self._sums(1, n.value)[0])
        self.op.adjoint_cl(scratch, x_tmp)
        _lib.call("dfm_fista_sums_partial", _lib.ptr(x_tmp), _lib.ptr(self.support_gpu), self.npix, None, None, 0, 0,
                  x_tmp.numel(), _lib.ptr(self._part), ctypes.byref(n), self.op._stream())
        at_u = float(self._sums(2, n.value)[0])
        if u_r > 0 and at_u > 0:
            self.adjoint_scale = at_u / u_r
        self._ew(0, D, a=x_tmp)

    def run(self, x0, b, niter, weights=None, gain=None, background=None, verbose=0, diagnostics_interval=10,
            state=None, callback=None):
        """niter iterations (fista_poisson.py:414-617). x0: a torch tensor on the operator's device or a NumPy array,
        either updated in place and returned. b, weights, gain, background: as SPDHGPoisson.run. state: the
        solver.state of an earlier run ({"it", "t", "theta", "y", "wls_until"}), x0 then that run's result.
        iter_stats per iteration: f, obj (with lam sum(x)), phase, theta, backtracks, local_metric, restart,
        metric_refresh, time; final_stats: niter, final_f, final_obj, backtracks, metric_updates_on_failure,
        adjoint_scale, restarts, metric_setup_s, time."""
        t_start = time.perf_counter()
        op = self.op
        dev = op.device
        cshape, dshape = tuple(op.coeff_shape), tuple(op.data_shape)
        host_x0 = None
        if isinstance(x0, torch.Tensor):
            x = check_device(x0, cshape, "x0", dev)
        else:
            host_x0 = as_host(x0, cshape, "x0")
            x = torch.from_numpy(host_x0.copy()).to(dev)
        bh = b if isinstance(b, torch.Tensor) else torch.from_numpy(as_host(b, dshape, "b"))
        if tuple(bh.shape) != dshape:
            raise ValueError(f"b must have shape {dshape}, not {tuple(bh.shape)}")
        self.b = bh.to(dev, torch.float32).contiguous()
        self.w, self.pw = _periodic(weights, dshape, "weights", dev)
        self.g, self.pg = _periodic(gain, dshape, "gain", dev)
        self.bg, self.pbg = _periodic(background, dshape, "background", dev)
        if self.pg and not float(self.g.min().item()) > 0:
            raise ValueError("the gain must be positive everywhere")
        xn = torch.empty(cshape, dtype=torch.float32, device=dev)
        grad = torch.empty(cshape, dtype=torch.float32, device=dev)
        D = torch.empty(cshape, dtype=torch.float32, device=dev)
        Ax = torch.empty(dshape, dtype=torch.float32, device=dev)
        Axn = torch.empty(dshape, dtype=torch.float32, device=dev)
        Ay = torch.empty(dshape, dtype=torch.float32, device=dev)
        self._ew(2, x)
        zero_start = state is None and float(x.max().item()) == 0.0
        wls_until = 0
        if self.start == "wls" and self.n_start > 0:
            if zero_start:
                wls_until = self.n_start
            elif state is not None and state.get("wls_until", 0) > int(state["it"]):
                wls_until = int(state["wls_until"])
        self._mode = 1 if wls_until else 0
        if zero_start and self.start == "uniform":
            self._ones_forward(Axn, xn)
            n = ctypes.c_int32(0)
            _lib.call("dfm_fista_sums_partial", _lib.ptr(Axn), None, 0, _lib.ptr(self.b), _lib.ptr(self.w), self.pw, 1,
                      Axn.numel(), _lib.ptr(self._part), ctypes.byref(n), op._stream())
            sb, sa = self._sums(2, n.value)
            if sa > 0 and sb > 0:
                x.copy_(xn)
                _lib.call("dfm_scale", _lib.ptr(x), float(sb / sa), x.numel(), op._stream())
        it0, t_mom, theta = 0, 1.0, 1.0
        if state is not None:
            it0, t_mom, theta = int(state["it"]), float(state["t"]), float(state["theta"])
            y = torch.from_numpy(as_host(state["y"], cshape, "state['y']")).to(dev)
        else:
            y = x.clone()
        op.direct_cl(x, Ax)
        if state is not None:
            op.direct_cl(y, Ay)
        else:
            Ay.copy_(Ax)
        self._metric(Ax, Axn, xn, D)
        t_metric = time.perf_counter() - t_start
        f_x = self._data_term(Ax, False)
        F_x = f_x + self.lam * float(x.sum(dtype=torch.float64).item())
        self.iter_stats = []
        n_bt_total, n_restart, n_metric = 0, 0, 0
        it = it0
        for k in range(int(niter)):
            it = it0 + k + 1
            f_y = self._data_term(Ay, True)            # Ay now holds df/dmu at y
            op.adjoint_cl(Ay, grad)
            theta = min(1.0, 2.0 * theta)
            n_bt = 0
            local = False
            while True:
                self._ew(3, xn, a=y, c=grad, D=D, th=theta, lam=self.lam)
                op.direct_cl(xn, Axn)
                f_n = self._data_term(Axn, False)
                nn = ctypes.c_int32(0)
                _lib.call("dfm_fista_step_sums_partial", _lib.ptr(xn), _lib.ptr(y), _lib.ptr(grad), _lib.ptr(D),
                          xn.numel(), _lib.ptr(self._part), ctypes.byref(nn), op._stream())
                p = self._sums(3, nn.value)
                lin = float(p[0]) / self.adjoint_scale
                quad = float(p[1]) / (theta * self.adjoint_scale)
                xsum = float(p[2])
                if f_n <= f_y + lin + 0.5 * quad + 1e-6 * abs(f_y) or quad == 0.0:
                    break
                n_bt += 1
                if not local and self.refresh_on_failure:
                    op.direct_cl(y, Axn)
                    self._metric(Axn, Ay, xn, D)
                    local, theta = True, 1.0
                    n_metric += 1
                    continue
                theta *= 0.5
                if n_bt > self.max_backtracks:
                    raise RuntimeError(f"FISTAPoisson: no sufficient decrease after {n_bt} step reductions (iteration "
                                       f"{it}, mode {self._mode}: f(y) {f_y:.10e}, f(x_new) {f_n:.10e}, derivative "
                                       f"{lin:.4e}, quadratic term {quad:.4e}, theta {theta:.3g})")
            n_bt_total += n_bt
            F_n = f_n + self.lam * xsum
            restarted = F_n > F_x
            if restarted:
                t_mom = 1.0
                n_restart += 1
            t_new = 0.5 * (1.0 + np.sqrt(1.0 + 4.0 * t_mom * t_mom))
            beta = float(np.float32((t_mom - 1.0) / t_new))
            t_mom = t_new
            self._ew(4, y, a=xn, c=x, th=beta)
            self._ew(4, Ay, a=Axn, c=Ax, th=beta)
            x.copy_(xn)
            Ax, Axn = Axn, Ax
            f_x, F_x = f_n, F_n
            switch = self._mode == 1 and it >= wls_until
            if switch:
                self._mode = 0
                f_x = self._data_term(Ax, False)
                F_x = f_x + self.lam * xsum
            refreshed = (it in self.refresh or switch) and k + 1 < niter and self._mode == 0
            if refreshed:
                self._metric(Ax, Axn, xn, D)
                t_mom, theta = 1.0, 1.0
                y.copy_(x)
                Ay.copy_(Ax)
            elif it % 50 == 0:
                op.direct_cl(y, Ay)
            stats = {"iter": it, "f": f_n, "obj": F_n, "phase": "wls" if it <= wls_until else "poisson",
                     "theta": theta, "backtracks": n_bt, "local_metric": local, "restart": bool(restarted),
                     "metric_refresh": bool(refreshed), "time": time.perf_counter() - t_start}
            self.iter_stats.append(stats)
            stop = bool(callback(it, x if host_x0 is None else x.cpu().numpy(), stats)) if callback else False
            if verbose and (k == 0 or (k + 1) % diagnostics_interval == 0 or k + 1 == niter or stop):
                print(f"[iter {it:5d}] f={f_n:.8e} obj={F_n:.8e} theta={theta:.3g} backtracks={n_bt_total} "
                      f"metric updates={n_metric} restarts={n_restart} {time.perf_counter() - t_start:.1f} s",
                      flush=True)
            if stop:
                break
        done = len(self.iter_stats)
        self.state = {"it": it0 + done, "t": t_mom, "theta": theta, "y": y.cpu().numpy(), "wls_until": wls_until}
        self.final_stats = {"niter": done, "final_f": f_x, "final_obj": F_x, "backtracks": n_bt_total,
                            "metric_updates_on_failure": n_metric, "adjoint_scale": self.adjoint_scale,
                            "restarts": n_restart, "metric_setup_s": t_metric, "time": time.perf_counter() - t_start}
        if host_x0 is not None:
            host_x0[...] = x.cpu().numpy()
            return host_x0
        return x
