"""Host-core baseline of bench.py (test infrastructure): times the C/OpenMP FP32 restatement of one SPDHG subset step
(oracle/c/cpu_step.c) on the benchmark's full workload.

The sample's inputs (CSR pole-figure matrices, step sizes, dual state) are taken from a GPU solver set up on the same
sample, so the baseline times exactly the per-step work: dual update, CSR adjoint, backprojection + primal update,
forward projection of the next subset, CSR forward.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
import time
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "c" / "libcpu_step.so"


def build():
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB


def _load():
    if not LIB.exists():
        build()
    lib = ctypes.CDLL(str(LIB))
    P, I, F, L = ctypes.c_void_p, ctypes.c_int, ctypes.c_float, ctypes.c_int64
    lib.cpu_set_threads.argtypes = [I]
    lib.cpu_radon_forward.argtypes = [P, P, P, I, I, I, I, I, F]
    lib.cpu_radon_backward_update.argtypes = [P, P, I, I, I, I, I, P, P, P, F, F]
    lib.cpu_spmm_forward.argtypes = [P, P, P, P, P, P, I, I, I, I]
    lib.cpu_spmm_adjoint.argtypes = [P, P, P, P, P, P, I, I, I, I, I, F]
    lib.cpu_dual_update.argtypes = [P, P, P, P, F, P, I, L]
    return lib


def _p(a):
    return ctypes.c_void_p(a.ctypes.data)


def time_steps(solver, budget_s=15.0, max_steps=400, threads=None):
    """Time the C step on the state of a GPU solver after start() (its operator must hold one orientation batch).
    Returns (steps per second, steps timed, threads)."""
    import torch
    op = solver.op
    if len(op.batches) != 1:
        raise ValueError("the baseline sample must fit one orientation batch")
    lib = _load()
    threads = threads or len(os.sched_getaffinity(0))
    lib.cpu_set_threads(int(threads))
    K = op.K_stride
    npix = op.Nx * op.Ny
    host = lambda t: np.ascontiguousarray(t.detach().cpu().numpy())
    sb = op.sparse_batches[0]
    rpf, ck, vf = host(sb["row_ptr_f"]), host(sb["col_k"]).view(np.uint16), host(sb["val_f"])
    rpa, cj, va = host(sb["row_ptr_a"]), host(sb["col_j"]).view(np.uint16), host(sb["val_a"])
    x, z, T = host(solver.xk), host(solver.zk), host(solver.Tk)
    y, sig, b, D = host(solver.y), host(solver.sig), host(solver.b), host(solver.D)
    torch.cuda.synchronize()
    geo_f = [np.ascontiguousarray(op.geo_fwd_cpu[s]) for s in op.omega_subsets]
    geo_b = [np.ascontiguousarray(op.geo_bwd_cpu[s]) for s in op.omega_subsets]
    om = [np.ascontiguousarray(s.astype(np.int32)) for s in op.omega_subsets]
    rmax = max(len(s) for s in op.omega_subsets)
    sino = np.empty(rmax * op.My * K, np.float32)
    row = op.My * op.N_seg
    alpha = float(op.N_Omega / np.pi)
    order = np.random.default_rng(0).permutation(solver.M).tolist()
    order = order + order

    def step(m, mn):
        na, nn = len(om[m]), len(om[mn])
        lib.cpu_dual_update(_p(y), _p(D), _p(sig), _p(b), float(np.float32(solver.eps_counts)), _p(om[m]), na, row)
        lib.cpu_spmm_adjoint(_p(D), _p(rpa), _p(cj), _p(va), _p(sino), _p(om[m]), na, op.My, op.N_seg, K,
                             op.batches[0]["K_batch"], alpha)
        lib.cpu_radon_backward_update(_p(sino), _p(geo_b[m]), op.Nx, op.Ny, op.My, na, K, _p(x), _p(z), _p(T),
                                      float(solver._inv_p_data), float(solver.lam))
        lib.cpu_radon_forward(_p(x), _p(sino), _p(geo_f[mn]), op.Nx, op.Ny, op.My, nn, K, float(op.radon_scale))
        lib.cpu_spmm_forward(_p(sino), _p(rpf), _p(ck), _p(vf), _p(D), _p(om[mn]), nn, op.My, op.N_seg, K)

    t0 = time.perf_counter()
    step(order[0], order[1])
    one = time.perf_counter() - t0
    n = int(max(2, min(max_steps, budget_s / max(one, 1e-6))))
    t0 = time.perf_counter()
    for i in range(n):
        step(order[(i + 1) % solver.M], order[(i + 2) % solver.M])
    dt = time.perf_counter() - t0
    return n / dt, n, int(threads)
