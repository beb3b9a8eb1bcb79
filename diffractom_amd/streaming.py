"""Forward and adjoint products with the coefficients in host memory (optimization/streaming.py:43-309, the
reference's host-streamed FISTA passes), for coefficient arrays larger than the device holds: the device keeps the
data-sized arrays and two batch-sized staging slots, independent of the number of orientations.

A (K, Ny, Nx) float32 host array moves one orientation batch of the operator at a time: a CPU copy (several
threads) into one of two page-locked staging buffers, a DMA on a copy stream into a device slot, and the batch's
products on the operator's stream (layout transpose, projection, PF product); the next batch's upload runs while
the current batch is computed, and in the adjoint the previous batch's download while the next is computed.
Streams and events order the slots: a slot is refilled only after the compute stream is done with it, a pinned
buffer only after its DMA finished. The products are the operator's own kernels in the same order as direct_cl /
adjoint_cl, so the results equal them bit for bit.
"""
from __future__ import annotations

import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import torch

from . import _lib


class HostStreamer:
    """streamer = HostStreamer(op); streamer.forward(x_host, data); streamer.adjoint(data, out_host)."""

    def __init__(self, op, n_threads=None):
        self.op = op
        dev = op.device
        self.npix = op.Nx * op.Ny
        n = self.npix * max(b["K_batch"] for b in op.batches)
        self.pinned = [torch.empty(n, dtype=torch.float32, pin_memory=True) for _ in range(2)]
        self.slots = [torch.empty(n, dtype=torch.float32, device=dev) for _ in range(2)]
        self.copy_stream = torch.cuda.Stream(device=dev)          # default priority
        self.pool = ThreadPoolExecutor(n_threads or min(8, os.cpu_count() or 1))
        self._slot_free = [None, None]        # event: the compute stream is done with slot i
        self._dma_done = [None, None]         # event: the last DMA of pinned buffer i finished

    def _compute_stream(self):
        op = self.op
        return op.queue if op.queue is not None else torch.cuda.current_stream(op.device)

    def _pcopy(self, dst, src):
        """dst[:] = src (NumPy, flat) with the thread pool (NumPy releases the GIL for large copies)."""
        parts = max(1, min(self.pool._max_workers, src.size // (1 << 20)))
        edges = np.linspace(0, src.size, parts + 1).astype(np.int64)
        list(self.pool.map(lambda i: np.copyto(dst[edges[i]:edges[i + 1]], src[edges[i]:edges[i + 1]]),
                           range(parts)))

    def _flat(self, x, name):
        x = np.asarray(x) if not isinstance(x, np.ndarray) else x
        shape = tuple(self.op.coeff_shape)
        if x.shape != shape or x.dtype != np.float32 or not x.flags.c_contiguous:
            raise ValueError(f"{name} must be a C-contiguous float32 host array of shape {shape}")
        return x.reshape(-1)

    def _upload(self, src, i):
        """host rows src -> pinned[i] -> slots[i] (DMA on the copy stream after the slot is free)."""
        if self._dma_done[i] is not None:
            self._dma_done[i].synchronize()
        n = src.size
        self._pcopy(self.pinned[i].numpy()[:n], src)
        cs = self.copy_stream
        if self._slot_free[i] is not None:
            cs.wait_event(self._slot_free[i])
        with torch.cuda.stream(cs):
            self.slots[i][:n].copy_(self.pinned[i][:n], non_blocking=True)
        ev = torch.cuda.Event()
        ev.record(cs)
        self._dma_done[i] = ev
        return ev

    def forward(self, x_host, data):
        """data = A x for a host array x (K, Ny, Nx); data is a device tensor of the data shape."""
        op = self.op
        src = self._flat(x_host, "x_host")
        st = self._compute_stream()
        sc = op._scratch_buffers(op.N_Omega)
        tab = op._tables(None)
        nb = len(op.batches)
        rows = lambda ib: src[self.npix * op.batches[ib]["k_start"]:self.npix * op.batches[ib]["k_end"]]
        up = self._upload(rows(0), 0)
        for ib, b in enumerate(op.batches):
            i = ib % 2
            st.wait_event(up)
            ks = b["kstride"]
            with torch.cuda.stream(st):
                _lib.call("dfm_gather_k", _lib.ptr(self.slots[i]), _lib.ptr(sc["img"]), self.npix, 0, b["K_batch"],
                          ks, _lib.stream_handle(st))
                ev = torch.cuda.Event()
                ev.record(st)
                self._slot_free[i] = ev
                op._radon_forward(sc["img"], ks, 0, ks, sc["sino"], op.geo_fwd, op.N_Omega, tab)
                op._pf_forward(ib, sc["sino"], data, None, accumulate=ib > 0)
            if ib + 1 < nb:                                 # the next batch's copies overlap this batch's products
                up = self._upload(rows(ib + 1), (ib + 1) % 2)
        return data

    def adjoint(self, data, out_host):
        """out = A^T data into a host array out (K, Ny, Nx); data is a device tensor of the data shape."""
        op = self.op
        dst = self._flat(out_host, "out_host")
        st = self._compute_stream()
        sc = op._scratch_buffers(op.N_Omega)
        tab = op._tables(None)
        pending = None                     # (pinned index, DMA event, host rows) of the previous batch
        for ib, b in enumerate(op.batches):
            i = ib % 2
            ks = b["kstride"]
            with torch.cuda.stream(st):
                if self._slot_free[i] is not None:
                    st.wait_event(self._slot_free[i])       # the slot's last download has been issued after this
                op._pf_adjoint(ib, data, sc["sino"], None)
                op._radon_backward(sc["sino"], op.geo_bwd, tab, op.N_Omega, sc["img"], ks, 0, ks)
                _lib.call("dfm_scatter_k", _lib.ptr(sc["img"]), _lib.ptr(self.slots[i]), self.npix, 0, b["K_batch"],
                          ks, _lib.stream_handle(st))
                done = torch.cuda.Event()
                done.record(st)
            n = self.npix * b["K_batch"]
            if self._dma_done[i] is not None:
                self._dma_done[i].synchronize()             # pinned[i] is free (its previous batch was copied out)
            cs = self.copy_stream
            cs.wait_event(done)
            with torch.cuda.stream(cs):
                self.pinned[i][:n].copy_(self.slots[i][:n], non_blocking=True)
            ev = torch.cuda.Event()
            ev.record(cs)
            self._dma_done[i] = ev
            self._slot_free[i] = ev
            if pending is not None:
                self._finish(*pending)
            pending = (i, ev, dst[self.npix * b["k_start"]:self.npix * b["k_end"]])
        if pending is not None:
            self._finish(*pending)
        return out_host

    def _finish(self, i, ev, rows):
        ev.synchronize()
        self._pcopy(rows, self.pinned[i].numpy()[:rows.size])
        self._dma_done[i] = None

    def release(self):
        self.pinned = self.slots = None
        self.pool.shutdown(wait=False)
