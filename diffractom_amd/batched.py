"""Many slices in one SPDHG run: slices that share the detector geometry but differ in orientation grid (K_i) and
support mask advance together, each kernel of a step launched once for all of them.

* Orientations of all slices are concatenated along the channel axis of the orientation-fastest arrays (slice i's
  block at columns off_i .. off_i + kstride_i, 64-aligned), so the projector kernels (which treat every channel
  independently) cover every slice in one launch; a slice's mask enters through its step size T (0 off its support).
* The PF products of the slices (each its own dense PF on the matrix cores) run as one grouped launch
  (dfm_mfma_grouped: a table maps every workgroup to its slice's product).
* Data-sized arrays are stacked per slice (n_slices x N_Omega rows), so the dual update of a subset is one launch
  over the rows of all slices.
Every slice follows the same subset order (one random stream, seed), so a batched run gives each slice the iterates
of SPDHGPoisson(op_i, subsets, seed=seed) run alone (spdhg_poisson.py:298-751), up to float32 summation order.
"""
from __future__ import annotations

import ctypes
import time

import numpy as np
import torch

from . import _lib
from .spdhg import SPDHGPoisson


class _Desc(ctypes.Structure):
    """dfm_mfma_desc of the C ABI."""
    _fields_ = [("src", ctypes.c_void_p), ("hi", ctypes.c_void_p), ("lo", ctypes.c_void_p),
                ("omegas", ctypes.c_void_p), ("out", ctypes.c_void_p),
                ("n_ang", ctypes.c_int32), ("my", ctypes.c_int32), ("cp", ctypes.c_int32), ("sks", ctypes.c_int32),
                ("kb", ctypes.c_int32), ("n_tiles", ctypes.c_int32), ("n_chunks", ctypes.c_int32),
                ("mt_n", ctypes.c_int32), ("n_yg", ctypes.c_int32), ("n_tg", ctypes.c_int32),
                ("ksplit", ctypes.c_int32), ("ch_per", ctypes.c_int32), ("direct", ctypes.c_int32),
                ("accumulate", ctypes.c_int32), ("alpha", ctypes.c_float), ("ncols", ctypes.c_int32)]


assert ctypes.sizeof(_Desc) == 104


class _BlkDesc(ctypes.Structure):
    """dfm_pfblk_desc of the C ABI."""
    _fields_ = [("src", ctypes.c_void_p), ("ptr", ctypes.c_void_p), ("pos", ctypes.c_void_p),
                ("val", ctypes.c_void_p), ("omegas", ctypes.c_void_p), ("ring_inv", ctypes.c_void_p),
                ("out", ctypes.c_void_p),
                ("n_ang", ctypes.c_int32), ("my", ctypes.c_int32), ("cp", ctypes.c_int32), ("sks", ctypes.c_int32),
                ("kb", ctypes.c_int32), ("P", ctypes.c_int32), ("n_kb", ctypes.c_int32), ("n_jb", ctypes.c_int32),
                ("n_cb", ctypes.c_int32), ("n_ch", ctypes.c_int32), ("mt_n", ctypes.c_int32),
                ("n_yg", ctypes.c_int32), ("n_tg", ctypes.c_int32), ("ksplit", ctypes.c_int32),
                ("ch_per", ctypes.c_int32), ("direct", ctypes.c_int32), ("accumulate", ctypes.c_int32),
                ("alpha", ctypes.c_float), ("ncols", ctypes.c_int32), ("blk_local", ctypes.c_int32)]


def _y_tiling(my):
    tiles = -(-my // 16)
    n_yg = -(-tiles // 8)
    return -(-tiles // n_yg), n_yg


class BatchedSPDHGPoisson:
    """SPDHGPoisson over many slices at once. operators: SinglePhaseForwardOperator of the slices (same geometry,
    PF on the matrix cores: dense fragments or blocks of the non-zeros, one 16-bit pf_precision for all, one
    orientation batch each); supports: one support option per slice (as SPDHGPoisson's `support`), or one for
    all."""

    def __init__(self, operators, subsets=8, support="fov", eps_counts=1e-3, rho="auto", seed=0, variant=2,
                 graphs=False):
        ops = list(operators)
        if not ops:
            raise ValueError("no slices")
        o0 = ops[0]
        if any(getattr(op, "coefficient_layout", "dense") == "local" for op in ops):
            raise NotImplementedError("BatchedSPDHGPoisson batches operators of the (K, Ny, Nx) coefficient layout")
        for op in ops:
            same = (op.Nx, op.Ny, op.My, op.N_Omega, op.N_seg) == (o0.Nx, o0.Ny, o0.My, o0.N_Omega, o0.N_seg) and \
                np.array_equal(op.angles, o0.angles) and op.cor_offset == o0.cor_offset
            if not same:
                raise ValueError("the slices of a batch must share the detector geometry (cfg)")
            if (op.mfma_batches is None and op.blk_batches is None) or len(op.batches) != 1:
                raise ValueError("batched slices need the PF on the matrix cores (pf_precision 'bf16', "
                                 "'bf16_residual' or 'fp16') in one orientation batch")
            if op.pf_precision != o0.pf_precision or (op.blk_batches is None) != (o0.blk_batches is None):
                raise ValueError("the slices of a batch must share pf_precision and the PF storage")
        supports = support if isinstance(support, (list, tuple)) else [support] * len(ops)
        if len(supports) != len(ops):
            raise ValueError("one support per slice")
        self.ops = ops
        self.solvers = [SPDHGPoisson(op, subsets=subsets, support=sp, eps_counts=eps_counts, rho=rho, seed=seed)
                        for op, sp in zip(ops, supports)]
        self.subsets = self.solvers[0].subsets
        self.M = len(self.subsets)
        self.rng = np.random.default_rng(seed)
        self._last = None
        self.variant = int(variant)
        self.residual = o0.pf_precision == "bf16_residual"
        self.blocks = o0.blk_batches is not None
        offs = np.cumsum([0] + [op.K_stride for op in ops])
        self.offsets = [int(v) for v in offs[:-1]]
        self.Kc = int(offs[-1])
        self.iter_stats = []
        # graphs: advance() replays each (subset, next subset) step from a captured HIP graph after running it
        # eagerly once and capturing it the second time (the launches of a step: dual update, grouped PF adjoint,
        # backprojection, forward projection, grouped PF forward; no host work, no allocation, no memset)
        self.graphs = bool(graphs)
        self._graph_cache, self._graph_seen = {}, set()

    # ---------------------------------------------------------------- order (spdhg_poisson.py:363-370)
    _epoch_order = SPDHGPoisson._epoch_order

    # ---------------------------------------------------------------- set-up
    def start(self, x0s, bs, weights=None):
        """Per-slice set-up (SPDHGPoisson.start: rho, S, T, support, x0), then the slices' arrays packed into the
        batch layout. x0s: one start per slice (NumPy or torch, as SPDHGPoisson.run's x0), or None: zeros (then
        finish() returns new device arrays)."""
        t0 = time.perf_counter()
        ops, o0 = self.ops, self.ops[0]
        dev = o0.device
        n = len(ops)
        npix = o0.Nx * o0.Ny
        R, My, CP = o0.N_Omega, o0.My, o0.N_seg
        self.xk = torch.zeros(npix * self.Kc, dtype=torch.float32, device=dev)
        self.zk = torch.zeros_like(self.xk)
        self.Tk = torch.zeros_like(self.xk)
        dshape = (n * R, My, CP)
        self.y = torch.empty(dshape, dtype=torch.float32, device=dev)
        self.sig = torch.empty_like(self.y)
        self.b = torch.empty_like(self.y)
        self.D = torch.zeros_like(self.y)
        self._x_out = []
        pw = None
        for i, (sol, op) in enumerate(zip(self.solvers, ops)):
            given = x0s is not None and x0s[i] is not None
            x0 = x0s[i] if given else torch.zeros(op.coeff_shape, dtype=torch.float32, device=dev)
            sol.start(x0, bs[i], weights=weights)
            xv = self.xk.view(npix, self.Kc)
            sl = slice(self.offsets[i], self.offsets[i] + op.K_stride)
            xv[:, sl] = sol.xk.view(npix, op.K_stride)
            self.zk.view(npix, self.Kc)[:, sl] = sol.zk.view(npix, op.K_stride)
            self.Tk.view(npix, self.Kc)[:, sl] = sol.Tk.view(npix, op.K_stride)
            self.y[i * R:(i + 1) * R] = sol.y
            self.sig[i * R:(i + 1) * R] = sol.sig
            self.b[i * R:(i + 1) * R] = sol.b
            if pw is None:
                self.w, self.pw = sol.w, sol.pw
                self.g, self.pg = sol.g, sol.pg
                self.bg, self.pbg = sol.bg, sol.pbg
                self.eps = sol.eps_counts
                self._inv_p = sol._inv_p_data
                pw = 1
            self._x_out.append((sol._x_dev if given else None, sol._x_on_gpu, sol._host_x))
            sol._x_dev = None
            self.rho_used = getattr(self, "rho_used", []) + [sol.rho_used]
            sol.xk = sol.zk = sol.Tk = sol.D = sol.y = sol.sig = None
        torch.cuda.synchronize(dev)
        # subset row lists of the stacked data, descriptors of the grouped PF products
        self._rows = [torch.from_numpy(np.concatenate([i * R + s for i in range(n)]).astype(np.int32)).to(dev)
                      for s in self.subsets]
        self._rows_all = torch.from_numpy(np.arange(n * R, dtype=np.int32)).to(dev)
        n_max = max(len(s) for s in self.subsets)
        self.sino = torch.empty(max(n_max, R) * My * self.Kc, dtype=torch.float32, device=dev)
        self._groups = {}
        self._graph_cache.clear()
        self._graph_seen.clear()
        self._seq = self._epoch_sequence()
        self._pos = 0
        self._forward(self._seq[0])
        self.t_setup = time.perf_counter() - t0

    def _epoch_sequence(self):
        return self._epoch_order()

    def _group(self, forward, m):
        """(descriptors, workgroup map, n_wg) of the grouped PF product over all slices for subset m (None: every
        angle)."""
        key = (forward, m)
        if key in self._groups:
            return self._groups[key]
        o0 = self.ops[0]
        R, My, CP = o0.N_Omega, o0.My, o0.N_seg
        n_ang = R if m is None else len(self.subsets[m])
        mt_n, n_yg = _y_tiling(My)
        if self.blocks:
            return self._blk_group(forward, m, key, n_ang, mt_n, n_yg)
        descs = (_Desc * len(self.ops))()
        wg = []
        for i, op in enumerate(self.ops):
            mb = op.mfma_batches[0]
            b = op.batches[0]
            d = descs[i]
            om = None if m is None else op._sub_omegas[m]
            d.omegas = om.data_ptr() if om is not None else None
            d.n_ang, d.my, d.cp, d.sks = n_ang, My, CP, self.Kc
            d.mt_n, d.n_yg, d.ksplit = mt_n, n_yg, 1
            d.direct, d.accumulate = 1, 0
            sino_i = self.sino.data_ptr() + 4 * self.offsets[i]
            data_i = self.D.data_ptr() + 4 * i * R * My * CP
            key2 = "fwd" if forward else "adj"
            d.hi = mb[key2 + "_hi"].data_ptr()
            d.lo = mb[key2 + "_lo"].data_ptr() if mb[key2 + "_lo"] is not None else None
            if forward:
                d.src, d.out = sino_i, data_i
                d.kb, d.n_tiles, d.n_chunks = b["K_batch"], mb["n_nt"], mb["n_kc"]
                d.ch_per, d.alpha, d.ncols = mb["n_kc"], 1.0, 0
            else:
                d.src, d.out = data_i, sino_i
                d.kb, d.n_tiles, d.n_chunks = b["kstride"], mb["n_kt"], mb["n_jc"]
                d.ch_per, d.alpha, d.ncols = mb["n_jc"], float(R / np.pi), b["kstride"]
            d.n_tg = -(-d.n_tiles // (4 * self.variant))
            n_wg = n_ang * n_yg * d.n_tg
            wg.append(np.stack([np.full(n_wg, i, np.int32), np.arange(n_wg, dtype=np.int32)], 1))
        wg = np.concatenate(wg)
        buf = torch.frombuffer(bytearray(bytes(descs)), dtype=torch.uint8).to(o0.device)
        wgm = torch.from_numpy(np.ascontiguousarray(wg)).to(o0.device)
        self._groups[key] = (buf, wgm, len(wg))
        return self._groups[key]

    def _blk_group(self, forward, m, key, n_ang, mt_n, n_yg):
        """The grouped block-sparse products (dfm_pfblk_grouped): one descriptor per slice, no k split (every
        slice's forward writes its data rows directly)."""
        o0 = self.ops[0]
        R, My, CP = o0.N_Omega, o0.My, o0.N_seg
        descs = (_BlkDesc * len(self.ops))()
        assert ctypes.sizeof(_BlkDesc) == _lib.load().dfm_pfblk_desc_size()
        wg = []
        for i, op in enumerate(self.ops):
            blk, b = op.blk_batches[0], op.batches[0]
            d = descs[i]
            om = None if m is None else op._sub_omegas[m]
            d.omegas = om.data_ptr() if om is not None else None
            d.ptr, d.pos, d.val = blk["ptr"].data_ptr(), blk["pos"].data_ptr(), blk["val"].data_ptr()
            d.ring_inv = blk["ring_inv"].data_ptr() if blk["ring_inv"] is not None else None
            d.n_ang, d.my, d.cp, d.sks, d.P = n_ang, My, CP, self.Kc, op.N_peaks
            d.n_kb, d.n_jb = blk["n_kb"], blk["n_jb"]
            d.mt_n, d.n_yg, d.ksplit = mt_n, n_yg, 1
            d.direct, d.accumulate = 1, 0
            sino_i = self.sino.data_ptr() + 4 * self.offsets[i]
            data_i = self.D.data_ptr() + 4 * i * R * My * CP
            if forward:
                d.src, d.out, d.kb = sino_i, data_i, b["K_batch"]
                d.n_cb, d.n_ch = blk["n_jb"], blk["n_kb"]
                d.alpha, d.ncols = 1.0, 0
            else:
                d.src, d.out, d.kb = data_i, sino_i, b["kstride"]
                d.n_cb, d.n_ch = -(-b["kstride"] // 32), blk["n_jb"]
                d.alpha, d.ncols = float(R / np.pi), b["kstride"]
            d.ch_per = d.n_ch
            d.n_tg = -(-d.n_cb // 4)
            n_wg = n_ang * n_yg * d.n_tg
            wg.append(np.stack([np.full(n_wg, i, np.int32), np.arange(n_wg, dtype=np.int32)], 1))
        wg = np.concatenate(wg)
        buf = torch.frombuffer(bytearray(bytes(descs)), dtype=torch.uint8).to(o0.device)
        wgm = torch.from_numpy(np.ascontiguousarray(wg)).to(o0.device)
        self._groups[key] = (buf, wgm, len(wg))
        return self._groups[key]

    def autotune(self, reps=3):
        """Time the projector variants of the first slice's operator on this batch's buffers (all slices'
        orientations at once; subset 0) and keep the fastest, as SinglePhaseForwardOperator.autotune does for one
        slice; the backprojection is timed on copies of x and z, so the solver state is untouched. Needs start()."""
        from .operator import event_timer
        o0 = self.ops[0]
        xc, zc = self.xk.clone(), self.zk.clone()
        self.tuning = o0.tune_projector(0, self.xk, self.Kc, 0, self.Kc, self.sino, xc, zc, self.Tk,
                                        event_timer(o0.device, reps))
        del xc, zc
        return self.tuning

    def _pf(self, forward, m):
        buf, wgm, n_wg = self._group(forward, m)
        if self.blocks:
            # + the slices' common y-tile count in bits 8-11 (a compiled form where there is one; same results)
            mt_n, _ = _y_tiling(self.ops[0].My)
            _lib.call("dfm_pfblk_grouped", int(forward), _lib.ptr(buf), _lib.ptr(wgm), n_wg,
                      (self.ops[0]._blk_mode("fwd") & 15) | (mt_n << 8), self.ops[0]._stream())
            return
        _lib.call("dfm_mfma_grouped", int(forward), _lib.ptr(buf), _lib.ptr(wgm), n_wg, int(self.residual),
                  self.variant, self.ops[0]._stream())

    def _forward(self, m):
        """D (rows of subset m of every slice, or every row) = A x: one projector launch, one grouped PF launch."""
        o0 = self.ops[0]
        n_ang, _, geo_f, _ = o0._angle_set(m)
        o0._radon_forward(self.xk, self.Kc, 0, self.Kc, self.sino, geo_f, n_ang, o0._tables(m))
        self._pf(True, m)

    def step(self, next_block=None, full_forward=False, block=None):
        """The step of subset `block` (default: the current one of the epoch order), then the forward of
        next_block or, with full_forward, of every angle."""
        o0 = self.ops[0]
        blk = self._seq[self._pos] if block is None else block
        row = o0.My * o0.N_seg
        rows = self._rows[blk]
        _lib.call("dfm_dual_update_sub", _lib.ptr(self.y), _lib.ptr(self.D), _lib.ptr(self.sig), _lib.ptr(self.b),
                  _lib.ptr(self.w), self.pw, _lib.ptr(self.g), self.pg, _lib.ptr(self.bg), self.pbg,
                  float(self.eps), _lib.ptr(rows), rows.numel(), row, o0._stream())
        self._pf(False, blk)
        n_ang, _, _, geo_b = o0._angle_set(blk)
        o0._radon_backward(self.sino, geo_b, o0._tables(blk), n_ang, None, self.Kc, 0, self.Kc, x=self.xk,
                           z=self.zk, T=self.Tk, inv_p=self._inv_p, lam=0.0)
        if full_forward:
            self._forward(None)
        elif next_block is not None:
            self._forward(next_block)

    def _run_step(self, nb):
        """One step followed by the forward of subset nb (the full forward when nb repeats the current subset),
        from a captured graph when graphs are on."""
        blk = self._seq[self._pos]
        full = nb == blk
        if not self.graphs or full or self.ops[0].queue is not None:
            self.step(nb, full_forward=full)
            return
        key = (blk, nb)
        g = self._graph_cache.get(key)
        if g is None and key in self._graph_seen:
            g = self._capture(blk, nb)
        if g is None:
            self._graph_seen.add(key)
            self.step(nb)
        else:
            g.replay()

    def _capture(self, blk, nb):
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            self.step(nb, block=blk)
        self._graph_cache[(blk, nb)] = g
        return g

    def capture_all(self):
        """Capture the step of every ordered pair of distinct subsets (graphs on; after start()): the descriptors
        and tables of every subset are made first, eagerly, so nothing is allocated while capturing. Returns the
        number of graphs."""
        if not self.graphs:
            raise ValueError("capture_all needs graphs=True")
        o0 = self.ops[0]
        for m in range(self.M):
            self._group(True, m)
            self._group(False, m)
            o0._tables(m)
        torch.cuda.synchronize(o0.device)
        for a in range(self.M):
            for b in range(self.M):
                if a != b and (a, b) not in self._graph_cache:
                    self._capture(a, b)
        return len(self._graph_cache)

    def advance(self, n_steps):
        for _ in range(int(n_steps)):
            if self._pos + 1 < len(self._seq):
                self._run_step(self._seq[self._pos + 1])
                self._pos += 1
            else:
                nseq = self._epoch_sequence()
                self._run_step(nseq[0])
                self._seq, self._pos = nseq, 0

    def deviances(self):
        """The Poisson deviance of every slice at the current D (which must hold A x on every row)."""
        out = []
        o0 = self.ops[0]
        R = o0.N_Omega
        for i, sol in enumerate(self.solvers):
            out.append(sol._deviance(self.D[i * R:(i + 1) * R]))
        return out

    def run(self, x0s, bs, n_epochs, weights=None, diagnostics_interval=1):
        """n_epochs epochs of every slice (the control flow of SPDHGPoisson.run); returns the slices' x. iter_stats
        per epoch hold every slice's objective where it was evaluated."""
        self.start(x0s, bs, weights)
        self.iter_stats = []
        seq = self._seq
        for e in range(n_epochs):
            last = e + 1 == n_epochs
            want_f = (e + 1) % max(1, diagnostics_interval) == 0 or last
            nseq = None
            f = None
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
                if (end and want_f) or nb is None or nb == blk:
                    self.step(None, full_forward=(end and want_f) or nb is not None)
                    if end and want_f:
                        f = self.deviances()
                else:
                    self.step(nb)
            if nseq is not None:
                seq = nseq
                self._seq = seq
            self.iter_stats.append({"epoch": e + 1, "f": f})
        return self.finish()

    def finish(self):
        """The slices' x in their own layout (NumPy or torch, as given to start)."""
        o0 = self.ops[0]
        npix = o0.Nx * o0.Ny
        outs = []
        for i, op in enumerate(self.ops):
            x_dev, on_gpu, host_x = self._x_out[i]
            if x_dev is None:
                x_dev = torch.empty(op.coeff_shape, dtype=torch.float32, device=o0.device)
            _lib.call("dfm_scatter_k", ctypes.c_void_p(self.xk.data_ptr() + 4 * self.offsets[i]), _lib.ptr(x_dev),
                      npix, 0, op.K, self.Kc, op._stream())
            if not on_gpu:
                torch.cuda.synchronize(o0.device)
                host_x[...] = x_dev.cpu().numpy()
                outs.append(host_x)
            else:
                outs.append(x_dev)
        torch.cuda.synchronize(o0.device)
        return outs
