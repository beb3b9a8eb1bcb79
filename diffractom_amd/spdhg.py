"""SPDHGPoisson on MI355X: stochastic primal-dual hybrid gradient over rotation-angle subsets for the Poisson
deviance, with the reference's signature and iteration (optimization/spdhg_poisson.py:298-751).

Iteration, for the next subset m of a random permutation per epoch (p_m = 1/M):
    y_m' = prox_{S F_m*}(y_m + S A_m x)            (float64 arithmetic, dfm_dual_update_sub)
    dz = A_m^T (y_m' - y_m); z += dz; x = max(0, x - T (z + dz/p_m + lam))   (fused into the backprojection)
    D_{m_next} = A_{m_next} x                       (forward of the next subset, into the same data-sized array)
with S = gamma/(A 1_S), T = min_m p_m/(gamma A_m^T 1_counted), gamma = sqrt(rho), rho = 3e4/mean(b)^2 ("auto").
With tv > 0 the total variation of every orientation image is one more block (probability p_tv, default 1/(M+1);
the subsets (1 - p_tv)/M each): its dual q per pixel is projected onto |q| <= tv after a step
0.99 p_tv / (4 max (T_j + T_j')), and x is updated with D^T dq (spdhg_poisson.py:108-193, 430-502).

MI355X design: x, z and T stay on the GPU for the whole run in the orientation-fastest layout (Npix, K) the Radon
kernels read, so a step is the subset's CSR adjoint, one backprojection kernel that also applies the primal update,
and the next subset's forward projection and CSR product: no gather/scatter passes and no dz array.
"""
from __future__ import annotations

import ctypes
import time

import numpy as np
import torch

from . import _lib
from .operator import as_host, check_device

RHO_AUTO = 3e4            # pdhg_poisson.py:33
NAMES = ("x", "z", "T")
TV_NAMES = ("qx", "qy")
TV_SAFETY = 0.99          # spdhg_poisson.py:277: the TV dual step is this fraction of its bound
DFM_MAX_PARTIALS = 2048


def _periodic(a, dshape, name, device):
    """(device tensor, period) of a per-segment, per-(translation, segment) or per-point array
    (fista_poisson.py:287-308); None -> (None, 0)."""
    if a is None:
        return None, 0
    n_seg, n_ys, n_all = dshape[-1], dshape[-2] * dshape[-1], int(np.prod(dshape))
    if isinstance(a, torch.Tensor):
        if a.dtype != torch.float32 or not a.is_contiguous() or a.numel() not in (n_seg, n_ys, n_all):
            raise ValueError(f"{name} must be C-contiguous float32 with N_seg = {n_seg}, My * N_seg = {n_ys} or "
                             f"{n_all} elements; got shape {tuple(a.shape)}, {a.dtype}")
        return a.reshape(-1).to(device), int(a.numel())
    h = np.ascontiguousarray(a, dtype=np.float32)
    if h.size == 1:
        h = np.full(n_seg, float(h.ravel()[0]), np.float32)
    if h.size not in (n_seg, n_ys, n_all):
        raise ValueError(f"{name} must have N_seg = {n_seg} elements, My * N_seg = {n_ys} or the data shape "
                         f"{tuple(dshape)}; got shape {h.shape}")
    return torch.from_numpy(h.ravel().copy()).to(device), int(h.size)


def is_local(op):
    """True for an operator whose coefficients are a flat vector of patches (the local basis)."""
    return getattr(op, "coefficient_layout", "dense") == "local"


def support_to_device(op, support):
    """The support option of the solvers as a flat (Ny * Nx) uint8 device mask or None (optimization/prox.py:75-99):
    "fov" (the operator's field of view), None / False (no support constraint) or an (Ny, Nx) array. A local-basis
    operator holds its support itself (spdhg_poisson.py:754-814 runs without a mask): None."""
    if support is None or support is False or is_local(op):
        return None
    if isinstance(support, str):
        if support != "fov":
            raise ValueError(f"support must be 'fov', None or an (Ny, Nx) array, not {support!r}")
        mask = op.support_mask()
    else:
        mask = np.asarray(support)
    if mask.shape != (op.Ny, op.Nx):
        raise ValueError(f"support mask shape {mask.shape} != (Ny, Nx) = {(op.Ny, op.Nx)}")
    return torch.from_numpy(np.ascontiguousarray(mask, dtype=np.uint8).ravel()).to(op.device)


class SPDHGPoisson:
    """SPDHG over omega subsets for the Poisson deviance."""

    def __init__(self, operator, subsets=8, prox_kind="nonneg", lam=0.0, support="fov", eps_counts=1e-3, rho="auto",
                 seed=0, resident="auto", tv=0.0, tv_prob=None, tv_tree=None):
        if prox_kind not in ("nonneg", "nonneg_l1"):
            raise ValueError("SPDHGPoisson supports prox_kind 'nonneg' and 'nonneg_l1'")
        if not (resident in ("auto", "all", "none") or set(resident) <= set(NAMES + ("qx", "qy"))):
            raise ValueError(f"resident must be 'auto', 'all', 'none' or a list of names from {NAMES + ('qx', 'qy')}")
        if float(tv) < 0:
            raise ValueError("tv must be >= 0")
        if tv_tree is not None:
            raise NotImplementedError("tv_tree needs an operator with the local coefficient layout")
        self.op = operator
        self._local = is_local(operator)
        self.subsets = operator.set_omega_subsets(subsets)
        self.M = len(self.subsets)
        self.lam = float(lam) if prox_kind == "nonneg_l1" else 0.0
        self.eps_counts = float(eps_counts)
        self.rho = rho if rho == "auto" else float(rho)
        self.rng = np.random.default_rng(seed)
        self.resident = resident
        # block probabilities (spdhg_poisson.py:339-346): data subsets (1 - p_tv) / M each, the TV block p_tv
        self.tv = float(tv)
        self.tv_prob = (1.0 / (self.M + 1) if tv_prob is None else float(tv_prob)) if self.tv > 0 else 0.0
        if self.tv > 0 and not 0 < self.tv_prob < 1:
            raise ValueError("tv_prob must be in (0, 1)")
        self._n_tv = max(1, int(round(self.M * self.tv_prob / (1 - self.tv_prob)))) if self.tv > 0 else 0
        self._p_data = (1.0 - self.tv_prob) / self.M
        self._inv_p_data = float(self.M) if self.tv == 0 else 1.0 / self._p_data
        self._last = None
        self.support_mask = self._support(support)
        self.state = None
        # HIP graphs of the subset steps: the launches of a step are two graphs, ("update", m) = dual update,
        # adjoint and primal update of subset m, and ("forward", m') = the forward of the next subset m'; each is
        # captured the second time it occurs and replayed
        # from then on (every array of the step is allocated before; a reallocated operator scratch buffer drops the
        # captured graphs). Only on the current stream (operator queue None).
        self.graphs = False
        self._graph_cache = {}
        self._graph_seen = set()
        self._graph_gen = None

    def _support(self, support):
        return support_to_device(self.op, support)

    # ---------------------------------------------------------------- subset order
    def _epoch_order(self):
        """A random permutation that never starts with the last subset of the previous epoch
        (spdhg_poisson.py:363-370)."""
        perm = self.rng.permutation(self.M)
        if self.M > 1 and int(perm[0]) == self._last:
            perm[[0, 1]] = perm[[1, 0]]
        self._last = int(perm[-1])
        return [int(m) for m in perm]

    def _epoch_sequence(self):
        """The blocks of the next epoch (spdhg_poisson.py:372-381): the subsets and, with tv, the TV block (-1) at
        _n_tv random places (without tv exactly the random numbers of _epoch_order)."""
        order = self._epoch_order()
        if self.tv <= 0:
            return order
        n = len(order) + self._n_tv
        tv_at = set(int(i) for i in self.rng.choice(n, self._n_tv, replace=False))
        it = iter(order)
        return [-1 if i in tv_at else next(it) for i in range(n)]

    # ---------------------------------------------------------------- reductions
    def _partials(self):
        return self._part

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

    # ---------------------------------------------------------------- steps
    def _dual(self, m):
        op = self.op
        row = op.My * op.N_seg
        _lib.call("dfm_dual_update_sub", _lib.ptr(self.y), _lib.ptr(self.D), _lib.ptr(self.sig), _lib.ptr(self.b),
                  _lib.ptr(self.w), self.pw, _lib.ptr(self.g), self.pg, _lib.ptr(self.bg), self.pbg,
                  float(self.eps_counts), _lib.ptr(op._sub_omegas[m]), len(self.subsets[m]), row, op._stream())

    def _adjoint_update(self, m):
        """z += A_m^T D; x updated (all batches): the CSR adjoint into the sinogram scratch, then the fused
        backprojection + primal update on the batch's columns of x, z, T."""
        op = self.op
        if self._local:
            op.adjoint_update(self.D, m, self.xk, self.zk, self.Tk, self._inv_p_data, self.lam)
            return
        n_ang, _, _, geo = op._angle_set(m)
        tab = op._tables(m)
        sc = op._scratch_buffers(n_ang)
        for ib, bt in enumerate(op.batches):
            op._pf_adjoint(ib, self.D, sc["sino"], m)
            op.mark("pf_adjoint")
            op._radon_backward(sc["sino"], geo, tab, n_ang, None, op.K_stride, bt["koff"], bt["kstride"],
                               x=self.xk, z=self.zk, T=self.Tk, inv_p=self._inv_p_data, lam=self.lam)
            op.mark("radon_backward_spdhg")

    def _forward(self, m):
        """D (rows of subset m, or all rows for m None) = A x."""
        self.op.direct_k(self.xk, self.D, subset=m)

    def _tv_update(self):
        """The TV block (spdhg_poisson.py:462-475): q <- proj(q + s D x), dz = D^T dq, z += dz, x updated with
        1/p_tv."""
        op = self.op
        st = op._stream()
        if self._local:
            n = self.xk.numel()
            _lib.call("dfm_tv_dual_nbr", _lib.ptr(self.xk), _lib.ptr(self.Tk), _lib.ptr(self.qxk), _lib.ptr(self.qyk),
                      _lib.ptr(self._dqx), _lib.ptr(self._dqy), _lib.ptr(self._nbr[0]), _lib.ptr(self._nbr[1]),
                      float(self.tv), float(np.float32(TV_SAFETY * self.tv_prob)), n, st)
            _lib.call("dfm_tv_adjoint_update_nbr", _lib.ptr(self._dqx), _lib.ptr(self._dqy), _lib.ptr(self._nbr[2]),
                      _lib.ptr(self._nbr[3]), _lib.ptr(self.xk), _lib.ptr(self.zk), _lib.ptr(self.Tk),
                      float(np.float32(1.0 / self.tv_prob)), float(self.lam), n, st)
            return
        _lib.call("dfm_tv_dual_k", _lib.ptr(self.xk), _lib.ptr(self.Tk), _lib.ptr(self.qxk), _lib.ptr(self.qyk),
                  _lib.ptr(self._dqx), _lib.ptr(self._dqy), op.Nx, op.Ny, op.K_stride, float(self.tv),
                  float(np.float32(TV_SAFETY * self.tv_prob)), st)
        _lib.call("dfm_tv_adjoint_update_k", _lib.ptr(self._dqx), _lib.ptr(self._dqy), _lib.ptr(self.xk),
                  _lib.ptr(self.zk), _lib.ptr(self.Tk), op.Nx, op.Ny, op.K_stride, float(np.float32(1.0 / self.tv_prob)),
                  float(self.lam), st)

    def _tv_value(self):
        op = self.op
        n = ctypes.c_int32(0)
        if self._local:
            _lib.call("dfm_tv_value_nbr_partial", _lib.ptr(self.xk), _lib.ptr(self.Tk), _lib.ptr(self._nbr[0]),
                      _lib.ptr(self._nbr[1]), self.xk.numel(), _lib.ptr(self._part), ctypes.byref(n), op._stream())
            return float(self._part[: n.value].sum().item())
        _lib.call("dfm_tv_value_partial", _lib.ptr(self.xk), _lib.ptr(self.Tk), op.Nx, op.Ny, op.K_stride,
                  _lib.ptr(self._part), ctypes.byref(n), op._stream())
        return float(self._part[: n.value].sum().item())

    # ---------------------------------------------------------------- run
    def start(self, x0, b, weights=None, gain=None, background=None, state=None):
        """Set-up of a run (spdhg_poisson.py:529-662): inputs on the device, rho, the step sizes S and T, x on the
        support, the dual state, and D = A x on the rows of the first subset. The steps then run with advance()."""
        t_start = time.perf_counter()
        op = self.op
        dev = op.device
        cshape, dshape = tuple(op.coeff_shape), tuple(op.data_shape)
        # graphs captured in an earlier run hold that run's buffer addresses (y, sig, D, b, w, x, z, T are
        # allocated anew below): none of them may be replayed in this run
        self._graph_cache.clear()
        self._graph_seen.clear()
        self._graph_gen = None
        if self.tv > 0 and state is not None and state.get("z") is None:
            raise ValueError("continuing a TV run needs the state's z")
        self._x_on_gpu = isinstance(x0, torch.Tensor)
        self._host_x = None if self._x_on_gpu else as_host(x0, cshape, "x0")
        self._x_dev = check_device(x0, cshape, "x0", dev) if self._x_on_gpu else \
            torch.from_numpy(self._host_x).to(dev)
        if isinstance(b, torch.Tensor):
            self.b = check_device(b, dshape, "b", dev)
        else:
            self.b = torch.from_numpy(as_host(b, dshape, "b")).to(dev)
        dummy = torch.zeros(1, dtype=torch.float32, device=dev)
        self.w, self.pw = _periodic(weights, dshape, "weights", dev)
        self.g, self.pg = _periodic(gain, dshape, "gain", dev)
        self.bg, self.pbg = _periodic(background, dshape, "background", dev)
        self.w = self.w if self.w is not None else dummy
        self.g = self.g if self.g is not None else dummy
        self.bg = self.bg if self.bg is not None else dummy
        self._part = torch.empty(2 * DFM_MAX_PARTIALS, dtype=torch.float64, device=dev)
        # the resident layout: channel-fastest (Npix, K_stride), or the local basis's flat vector as (1, n_coef)
        npix, ks = (1, op.n_coef) if self._local else (op.Nx * op.Ny, op.K_stride)
        nk = npix * ks
        st = op._stream()
        supp = self.support_mask
        self.y = torch.zeros(dshape, dtype=torch.float32, device=dev)
        self.sig = torch.empty(dshape, dtype=torch.float32, device=dev)
        self.D = torch.empty(dshape, dtype=torch.float32, device=dev)
        nd = int(np.prod(dshape))

        rho = self.rho
        if rho == "auto":
            mean_b = self._mean_counted(self.b)
            rho = RHO_AUTO / mean_b ** 2 if mean_b > 0 else 1.0
        self.rho_used = float(rho)
        gamma = float(np.sqrt(rho))
        p = self._p_data

        # resident channel-fastest arrays
        self.xk = op.to_k(self._x_dev)
        self.zk = torch.zeros(nk, dtype=torch.float32, device=dev)
        self.Tk = torch.empty(nk, dtype=torch.float32, device=dev)
        self.resident_used = NAMES

        # S from the row sums of A 1_S (spdhg_poisson.py:621-631)
        if self._local:
            ones = torch.from_numpy(op.coefficient_mask()).to(dev)
        else:
            ones = torch.empty(nk, dtype=torch.float32, device=dev)
            _lib.call("dfm_support_ones_k", _lib.ptr(ones), _lib.ptr(supp), npix, ks, ks, st)
            self._zero_padding(ones)
        op.direct_k(ones, self.D)
        self.scales = {"row_sum_mean": self._mean_counted(self.D), "b_mean": self._mean_counted(self.b),
                       "gain_mean": 1.0 if self.pg == 0 else float(self.g.mean().item())}
        _lib.call("dfm_dual_steps", _lib.ptr(self.sig), _lib.ptr(self.D), _lib.ptr(self.w), self.pw, gamma, nd, st)
        # T = min_m p / (gamma A_m^T 1_counted) on the support (spdhg_poisson.py:632-643)
        _lib.call("dfm_set_mask_ones", _lib.ptr(self.D), _lib.ptr(self.w), self.pw, nd, st)
        _lib.call("dfm_tinit", _lib.ptr(self.Tk), npix, ks, st)
        colsum = ones
        for m in range(self.M):
            op.adjoint_k(self.D, colsum, subset=m)
            _lib.call("dfm_tmin", _lib.ptr(self.Tk), _lib.ptr(colsum), _lib.ptr(supp), npix, ks, ks,
                      float(np.float32(p / gamma)), st)
        _lib.call("dfm_tfinal", _lib.ptr(self.Tk), npix, ks, ks, st)
        self._zero_padding(self.Tk)
        del ones, colsum
        _lib.call("dfm_project_support_k", _lib.ptr(self.xk), _lib.ptr(supp), npix, ks, ks, st)
        self._zero_padding(self.xk)
        if state is not None:
            self.y.copy_(torch.from_numpy(as_host(state["y"], dshape, "state['y']")).to(dev))
            if state.get("z") is not None:
                zc = torch.from_numpy(as_host(state["z"], cshape, "state['z']")).to(dev)
                op.to_k(zc, self.zk)
            else:
                op.adjoint_k(self.y, self.zk)
        if self.tv > 0:
            self.qxk = torch.zeros(nk, dtype=torch.float32, device=dev)
            self.qyk = torch.zeros(nk, dtype=torch.float32, device=dev)
            for nm, qk in (("qx", self.qxk), ("qy", self.qyk)):
                if state is not None and state.get(nm) is not None:
                    op.to_k(torch.from_numpy(as_host(state[nm], cshape, f"state['{nm}']")).to(dev), qk)
            self._dqx = torch.empty(nk, dtype=torch.float32, device=dev)
            self._dqy = torch.empty(nk, dtype=torch.float32, device=dev)
            if self._local:
                self._nbr = [torch.from_numpy(a).to(dev) for a in op.patch_neighbours()]
            self.resident_used = NAMES + TV_NAMES
        self._ep0 = int(state["epoch"]) if state is not None else 0
        self._seq = self._epoch_sequence()
        self._pos = 0
        if self._seq[0] >= 0:
            op.zero_subset_rows(self.D, self._seq[0])
            self._forward(self._seq[0])
        self._t_start = t_start
        self.t_setup = time.perf_counter() - t_start

    def step(self, next_block=None, full_forward=False):
        """One step of the current block (spdhg_poisson.py:430-484 _pass): a data subset's dual update, adjoint and
        primal update, or the TV block (-1); then the forward of next_block when it is a data subset (or of every
        row with full_forward) into D."""
        blk = self._seq[self._pos]
        m_next = next_block if next_block is not None and next_block >= 0 else None
        if self.graphs and self.op.queue is None and not full_forward and m_next is not None and \
                self.op.marks is None:
            self._graph_run(("update", blk), lambda: self._step_launches(blk, None, False))
            self._graph_run(("forward", m_next), lambda: self._forward(m_next))
            return
        self._step_launches(blk, m_next, full_forward)

    def _graph_run(self, key, launches):
        """Run launches eagerly the first time key occurs, capture them the second time, replay them after."""
        gen = getattr(self.op, "scratch_generation", 0)
        if gen != self._graph_gen:
            self._graph_cache.clear()
            self._graph_seen.clear()
            self._graph_gen = gen
        g = self._graph_cache.get(key)
        if g is None and key in self._graph_seen:
            g = torch.cuda.CUDAGraph()
            with torch.cuda.graph(g):
                launches()
            if getattr(self.op, "scratch_generation", 0) != gen:    # a buffer was allocated while capturing
                self._graph_cache.clear()
                self._graph_seen.clear()
                self._graph_gen = getattr(self.op, "scratch_generation", 0)
                launches()
                return
            self._graph_cache[key] = g
        if g is None:
            self._graph_seen.add(key)
            launches()
        else:
            g.replay()

    def _step_launches(self, blk, m_next, full_forward):
        if blk >= 0:
            self._dual(blk)
            self.op.mark("dual_update")
            self._adjoint_update(blk)
        else:
            self._tv_update()
            self.op.mark("tv")
        if full_forward:
            self._forward(None)
        elif m_next is not None:
            self._forward(m_next)

    def advance(self, n_steps):
        """n subset steps in the epoch order (no objective evaluations): each step ends with the forward of the next
        subset, as in the middle of an epoch."""
        for _ in range(int(n_steps)):
            if self._pos + 1 < len(self._seq):
                nb = self._seq[self._pos + 1]
                self.step(nb, full_forward=nb >= 0 and nb == self._seq[self._pos])
                self._pos += 1
            else:
                nseq = self._epoch_sequence()
                nb = nseq[0]
                self.step(nb, full_forward=nb >= 0 and nb == self._seq[self._pos])
                self._seq, self._pos = nseq, 0

    def finish(self):
        """Write x back (and the state), free the run's arrays; returns x as run() does."""
        op = self.op
        dev = op.device
        cshape = tuple(op.coeff_shape)
        op.from_k(self.xk, self._x_dev)
        zc = torch.empty(cshape, dtype=torch.float32, device=dev)
        op.from_k(self.zk, zc)
        torch.cuda.synchronize(dev)
        self.state = {"epoch": self._ep0 + len(getattr(self, "iter_stats", [])), "z": zc.cpu().numpy(),
                      "y": self.y.cpu().numpy()}
        if self.tv > 0:
            for nm, qk in (("qx", self.qxk), ("qy", self.qyk)):
                op.from_k(qk, zc)
                self.state[nm] = zc.cpu().numpy()
            self.qxk = self.qyk = self._dqx = self._dqy = self._nbr = None
        if not self._x_on_gpu:
            self._host_x[...] = self._x_dev.cpu().numpy()
        out = self._x_dev if self._x_on_gpu else self._host_x
        self.b = self.w = self.g = self.bg = self._part = self.y = self.sig = self.D = None
        self.xk = self.zk = self.Tk = None
        return out

    def run(self, x0, b, n_epochs, weights=None, gain=None, background=None, verbose=0, diagnostics_interval=1,
            callback=None, state=None):
        """spdhg_poisson.py:513-751. x0: a torch tensor on the GPU (updated in place and returned) or a NumPy array
        (updated in place if C-contiguous float32, else a converted copy is returned). callback(epoch, x, stats)
        after every epoch; a true return stops the run. state: solver.state of an earlier run (z and y; without
        "z", z = A^T y is recomputed)."""
        op = self.op
        dev = op.device
        self.start(x0, b, weights, gain, background, state)
        self.iter_stats = []
        f = float("nan")
        seq = self._seq
        for e in range(n_epochs):
            ep = self._ep0 + e + 1
            last = e + 1 == n_epochs
            want_f = (e + 1) % max(1, diagnostics_interval) == 0 or last
            nseq = None
            for j, blk in enumerate(seq):
                self._pos = j
                end = j + 1 == len(seq)
                if not end:
                    nb = seq[j + 1]
                elif not last:
                    nseq = self._epoch_sequence()
                    nb = nseq[0]
                else:
                    nb = None
                adj = blk if blk >= 0 else None
                m_next = nb if nb is not None and nb >= 0 else None
                if (end and want_f) or nb is None or (m_next is not None and m_next == adj):
                    full = (end and want_f) or m_next is not None
                    self.step(None, full_forward=full)
                    if end and want_f:
                        f = self._deviance(self.D)
                        if self.tv > 0:
                            f += self.tv * self._tv_value()
                else:
                    self.step(nb)
            if nseq is not None:
                seq = nseq
                self._seq = seq
            torch.cuda.synchronize(dev)
            stats = {"epoch": ep, "f": f if want_f else None, "rho": self.rho_used,
                     "time": time.perf_counter() - self._t_start}
            self.iter_stats.append(stats)
            stop = False
            if callback is not None:
                op.from_k(self.xk, self._x_dev)
                if not self._x_on_gpu:
                    self._host_x[...] = self._x_dev.cpu().numpy()
                stop = bool(callback(ep, self._x_dev if self._x_on_gpu else self._host_x, stats))
            if verbose and want_f:
                print(f"[spdhg epoch {ep:4d}] f={f:.8e} {time.perf_counter() - self._t_start:.1f} s", flush=True)
            if stop:
                break
        out = self.finish()
        self.final_stats = {"epochs": len(self.iter_stats), "final_f": f, "setup_s": self.t_setup,
                            "rho": self.rho_used, "subsets": self.M, "resident": self.resident_used,
                            "time": time.perf_counter() - self._t_start, "tv": self.tv, "tv_prob": self.tv_prob,
                            **self.scales}
        return out

    def _zero_padding(self, a_k):
        """Padding channels (K_batch < kstride) of a channel-fastest array to 0."""
        op = self.op
        if self._local:
            return
        npix = op.Nx * op.Ny
        v = a_k.view(npix, op.K_stride)
        for bt in op.batches:
            if bt["kstride"] > bt["K_batch"]:
                v[:, bt["koff"] + bt["K_batch"]: bt["koff"] + bt["kstride"]] = 0.0
