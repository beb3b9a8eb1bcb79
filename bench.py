#!/usr/bin/env python3
"""SPDHG-Poisson subset steps per second on one MI355X, BF16 matrix-core PF products against the FP32 path
(workload "al_c1": Al, 180 omega x 103 translations x 180 eta x 5 rings, 99 x 99 pixels, K = 29713 orientations of
a uniform 3-degree cubic fundamental-zone grid, sigma 2 deg, 10 omega subsets, Poisson data at a mean of 1000 counts).

A step is one SPDHG subset iteration: the dual update of subset m, the PF adjoint and the backprojection fused with
the primal update, the forward projection and PF product of the next subset. The warm-up and timed steps continue
one run (the solver's own random subset order). Inputs and all solver arrays are resident in HBM before timing.

Four PF-product precisions run on the same problem: "fp32" (CSR SpMM; the tests pin it to the FP64 oracle within
2e-5) and three matrix-core paths on the PF stored as 32 x 32 blocks of its non-zeros: "bf16" (bf16 operands),
"bf16_residual" (plus the hi/lo residual correction) and "fp16" (fp16 operands with power-of-two scales per ring and
per row chunk), FP32 accumulation. The headline (--precision, default bf16_residual) reports its steps/s; every path
reports its forward, adjoint and SPDHG-iterate relative L2 difference to the fp32 path, all solving the same counts,
and its forward / adjoint error against the FP64 oracle on a sample.

Prints one JSON line (metric, value, roofline of the headline path's dominant kernel from HIP events over the timed
steps, cpu_baseline from the C/OpenMP restatement on the host cores, forward/adjoint GB/s, per-path figures).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
MFMA_BF16_PEAK_TFLOPS = 2500.0     # dense bf16 / fp16 matrix-core peak of one MI355X


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=60)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--config", default="al_c1", help="al_c1 (the headline line), al_c0, al_c3 (the multi-slice "
                    "batch: slices/s and the step-time percentiles), ti_c2 (hexagonal, 256^2 px, K 59,440; PF per "
                    "--c2-precision), "
                    "al_c4 (autograd loss + gradients, 8 slices), al_local (SPDHG on a local basis of the al_c1 "
                    "scan against the global operator on the same orientations), al_extras (parameter gradients, "
                    "Fourier-slice projector and host streaming on the al_c1 workload)")
    ap.add_argument("--slices", type=int, default=None, help="al_c3: number of slices (default 64)")
    ap.add_argument("--no-graphs", action="store_true", help="al_c3: eager launches instead of HIP graph replay")
    ap.add_argument("--compare-steps", type=int, default=300, help="the precisions' SPDHG iterates are compared "
                    "after this many steps of the run (the untimed steps after the timed ones)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=15.0, help="seconds of host-core baseline steps")
    ap.add_argument("--no-kernel-timing", action="store_true")
    ap.add_argument("--c2-precision", default="fp16", choices=["fp16", "bf16_residual", "fp32"],
                    help="ti_c2: the PF product (fp16 / bf16_residual: generated per call on the matrix cores; "
                         "fp32: stored CSR)")
    ap.add_argument("--precision", default="bf16_residual", choices=["bf16", "bf16_residual", "fp16", "fp32"],
                    help="PF product of the headline figure (the other precisions run too, for the comparison)")
    ap.add_argument("--only", action="store_true", help="run the headline precision only")
    ap.add_argument("--no-op-timing", action="store_true", help="skip the full forward / adjoint timing")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default kernel variants (no timing pass)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="end with a 1 GiB device-to-device copy (the reference kernel of the FETCH_SIZE / WRITE_SIZE "
                         "counter runs that give roofline.traffic)")
    ap.add_argument("--graphs", action="store_true", help="replay subset steps from captured HIP graphs (pays off on "
                    "small problems, where launches cost more than the kernels)")
    ap.add_argument("--no-compare", action="store_true", help="al_local: the local operator's steps only (no global "
                    "operator run, no parity figures; for profiling)")
    ap.add_argument("--solver", default="spdhg", choices=["spdhg", "osem", "pdhg", "fista"],
                    help="al_c0 / al_c1: the solver timed (osem: ordered-subsets EM subset iterations; pdhg: "
                         "Chambolle-Pock iterations over all angles; fista: SQS-metric FISTA iterations over all "
                         "angles; one precision)")
    return ap.parse_args()


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def build_problem(name, K=None, mean_counts=None, precision="fp32", counts=None):
    """Operator, Poisson counts and the truth of workload `name` (dense PF on the matrix cores for the bf16
    precisions, sparse CSR for fp32). counts: reuse these counts instead of drawing them from this operator's
    forward projection (so that every precision solves the same problem)."""
    from diffractom_amd import SinglePhaseForwardOperator
    from diffractom_amd.workloads import CONFIGS, make_grid, poisson_counts, voronoi_truth
    p = CONFIGS[name]
    grid, mat, cfg = make_grid(name, K=K)
    op = SinglePhaseForwardOperator(cfg, mat, grid, max_gb=1024.0, normalized=True, pf_mode="sparse",
                                    pf_precision=precision)
    xt = voronoi_truth(op.K, p["n"], p["grains"], p["radius"], p["seed"], p["coeff"])
    if counts is not None:
        return op, counts, xt, p
    ax = op.direct(xt)
    b, _ = poisson_counts(ax, mean_counts or p["mean_counts"], p["seed"])
    return op, b, xt, p


def timed_steps(solver, n_steps, per_kernel=True):
    """Run n_steps subset steps of the solver (advance()) between two HIP events; with per_kernel the solver's own
    launches record an event after every kernel class (op.marks), so the per-kernel durations come from the same
    timed steps. Returns (ms per step, {kernel class: average ms per step} or None)."""
    import torch
    op = solver.op
    op.marks = [] if per_kernel else None
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    solver.advance(n_steps)
    e1.record()
    torch.cuda.synchronize()
    marks, op.marks = op.marks, None
    ms = e0.elapsed_time(e1) / n_steps
    if not per_kernel:
        return ms, None
    tot = {}
    prev = e0
    for name, e in marks:
        tot[name] = tot.get(name, 0.0) + prev.elapsed_time(e)
        prev = e
    return ms, {k: v / n_steps for k, v in tot.items()}


def pf_nnz(op):
    """Non-zeros of the PF matrix (all angles, all orientations): the CSR's count, or the count the dense / fragment
    forms recorded when they were built."""
    return int(op.sparse_nnz)


def kernel_bytes(op, M):
    """Algorithmic HBM bytes per launch of each kernel class of one SPDHG step (subset of R/M angles): what the
    step must move whatever the storage form. The PF products: 6 bytes per non-zero of the subset's PF (one CSR
    form: a 4-byte value, a 2-byte index) + the sinogram + the data rows, also for the matrix-core path (its
    zero-padded fragment stream is reported beside it as `fragment_bytes`, not counted as algorithmic)."""
    npix = op.Nx * op.Ny
    K, R, My, CP = op.K, op.N_Omega, op.My, op.N_seg
    Rm = R / M
    sino = 4.0 * Rm * My * K
    rows = 4.0 * Rm * My * CP
    pf = 6.0 * pf_nnz(op) / M
    return {"radon_backward_spdhg": 20.0 * K * npix + sino, "radon_forward": 4.0 * K * npix + sino,
            "pf_adjoint": pf + rows + sino, "pf_forward": pf + sino + rows, "dual_update": 6.0 * rows}


def mfma_figures(op, M, kt):
    """The matrix-core PF products' fragment stream (bytes per launch) and MFMA utilisation: FLOPs the MFMAs issue
    per launch (2 Rm My_p CP_p K_p per product pass, the padded tile sizes; 3 passes with the residual) and the
    dense-algorithmic FLOPs 2 Rm My CP K, both over the kernel's HIP-event time against the dense BF16 peak."""
    if op.mfma_batches is None and op.blk_batches is None:
        return None
    R, My, CP, K = op.N_Omega, op.My, op.N_seg, op.K
    Rm = R / M
    out = {}
    for name, keys in (("pf_forward", ("fwd_hi", "fwd_lo")), ("pf_adjoint", ("adj_hi", "adj_lo"))):
        if name not in kt:
            continue
        passes = 3 if op.pf_precision.endswith("residual") else 1
        myp = -(-My // 16) * 16
        if op.blk_batches is not None:
            # blocks of non-zeros: (pos, value) per entry + the block pointers, the subset's share
            frag = sum(blk["pos"].numel() * 2 + blk["val"].numel() * blk["val"].element_size() +
                       blk["ptr"].numel() * 4 for blk in op.blk_batches) * Rm / R
            cpp = -(-CP // 32) * 32
            kp = sum(-(-b["K_batch"] // 32) * 32 for b in op.batches)
        else:
            frag = sum(mb[k].numel() * 2 for mb in op.mfma_batches for k in keys if mb[k] is not None) * Rm / R
            cpp = -(-CP // (16 if name == "pf_forward" else 32)) * (16 if name == "pf_forward" else 32)
            kp = sum(-(-b["K_batch"] // (32 if name == "pf_forward" else 16)) * (32 if name == "pf_forward" else 16)
                     for b in op.batches)
        issued = 2.0 * Rm * myp * cpp * kp * passes
        dense = 2.0 * Rm * My * CP * K
        t = kt[name] * 1e-3
        out[name] = {"fragment_bytes": frag, "fragment_GBps": round(frag / t / 1e9, 1),
                     "mfma_flops_issued": issued, "mfma_TFLOPs_issued": round(issued / t / 1e12, 2),
                     "mfma_frac_issued": round(issued / t / 1e12 / MFMA_BF16_PEAK_TFLOPS, 4),
                     "dense_flops": dense, "dense_TFLOPs": round(dense / t / 1e12, 2),
                     "peak_TFLOPs": MFMA_BF16_PEAK_TFLOPS}
    return out


def rel_l2(a, b):
    import torch
    a, b = a.double(), b.double()
    return float(torch.linalg.vector_norm(a - b) / torch.linalg.vector_norm(b))


SAMPLE_ROWS = (0, 45, 90, 135)      # angle rows of the FP64-oracle sample (both sweep classes)
SAMPLE_K0, SAMPLE_KN = 12288, 1024   # orientation block of the sample


def cpu_baseline_leg(args, op, b, M):
    """The host-side leg of the benchmark, the only part that runs code under oracle/ (as the CPU baseline and as
    the checker):
    (1) the C/OpenMP FP32 restatement of one SPDHG subset step (oracle/c/cpu_step.c: dual update, CSR adjoint,
        backprojection with the primal update, forward projection, CSR forward) timed on the full workload, all K
        orientations, on the state of a solver started on this FP32-CSR operator;
    (2) the FP64 oracle (oracle/model.py) on a sample of the workload: the forward rows of SAMPLE_ROWS for
        coefficients supported on the orientation block SAMPLE_K0 .. +SAMPLE_KN (seeded U(0, 1) on the support),
        and the adjoint of the counts on those rows restricted to the block. Every precision's GPU forward / adjoint
        of the same inputs is compared with these values (fp64_errors).
    Returns (cpu_baseline dict or None, oracle sample dict or None)."""
    import torch
    from diffractom_amd import SPDHGPoisson
    cpu = smp = None
    try:
        from oracle import cpu_baseline
        sol = SPDHGPoisson(op, subsets=M, seed=0)
        sol.start(np.zeros(op.coeff_shape, np.float32), b)
        torch.cuda.synchronize()
        rate, n, thr = cpu_baseline.time_steps(sol, budget_s=args.cpu_budget)
        sol.finish()
        cpu = {"value": round(rate, 4), "unit": "SPDHG steps/s", "cores": thr, "kind": "port",
               "sample": f"{n} consecutive SPDHG subset steps of the C/OpenMP FP32 CSR restatement on the full "
                         f"workload ({op.K} orientations, {op.N_Omega // M} angles per subset), {thr} threads"}
        log(f"[bench] cpu baseline: {rate:.4f} steps/s ({n} full-size steps, {thr} threads)")
    except Exception as exc:                                        # the GPU figures stand without it
        log(f"[bench] cpu baseline failed: {exc!r}")
    try:
        from oracle.model import OracleOperator
        from diffractom_amd.workloads import AL_RINGS
        t0 = time.perf_counter()
        ks = np.arange(SAMPLE_K0, min(SAMPLE_K0 + SAMPLE_KN, op.K))
        rows = np.array([r for r in SAMPLE_ROWS if r < op.N_Omega])
        orc = OracleOperator(op.cfg, (4.05, 4.05, 4.05, 90, 90, 90), "cubic", AL_RINGS, 0.2, op.grid_rot_cpu[ks],
                             float(op.sigma_cpu[0]), orthonormalize=False)
        rng = np.random.default_rng(11)
        xb = rng.random((len(ks), op.Ny, op.Nx)) * op.support_mask()[None]
        fwd = orc.direct_rows(xb, rows)
        orc.set_omega_subsets([rows])
        yb = np.zeros(op.data_shape)
        yb[rows] = np.asarray(b, np.float64)[rows]
        adj = orc.adjoint(yb, subset=0)
        t_orc = time.perf_counter() - t0
        smp = {"rows": rows, "ks": ks, "x": xb.astype(np.float32), "y_rows": yb[rows].astype(np.float32),
               "fwd": fwd, "adj": adj, "seconds": t_orc}
        log(f"[bench] FP64 oracle sample ({len(rows)} angle rows x {len(ks)} orientations): {t_orc:.1f} s")
    except Exception as exc:
        log(f"[bench] FP64 oracle sample failed: {exc!r}")
    return cpu, smp


def fp64_errors(op, smp):
    """Relative L2 errors of this operator's forward rows and block adjoint against the FP64 oracle sample."""
    import torch
    if smp is None:
        return None
    ks, rows = smp["ks"], smp["rows"]
    x = torch.zeros(op.coeff_shape, dtype=torch.float32, device="cuda")
    x[ks[0]:ks[-1] + 1] = torch.from_numpy(smp["x"]).cuda()
    fwd = op.direct(x)[torch.from_numpy(rows).cuda()].double().cpu().numpy()
    y = torch.zeros(op.data_shape, dtype=torch.float32, device="cuda")
    y[torch.from_numpy(rows).cuda()] = torch.from_numpy(smp["y_rows"]).cuda()
    adj = op.adjoint(y)[ks[0]:ks[-1] + 1].double().cpu().numpy()
    del x, y
    e = lambda a, r: float(np.linalg.norm(a - r) / np.linalg.norm(r))
    return {"forward": e(fwd, smp["fwd"]), "adjoint": e(adj, smp["adj"]),
            "sample": f"angle rows {list(map(int, rows))} x orientations {int(ks[0])}..{int(ks[-1])}"}


def run_path(args, precision, counts=None, leg=None):
    """One precision: set-up, warm-up, timed SPDHG steps, per-kernel times, full forward / adjoint times; returns
    the figures and the device results the precisions are compared on."""
    import torch
    from diffractom_amd import SPDHGPoisson
    t0 = time.perf_counter()
    op, b, xt, p = build_problem(args.config, precision=precision, counts=counts)
    M = p["subsets"]
    torch.cuda.synchronize()
    setup = time.perf_counter() - t0
    log(f"[bench] {precision}: K={op.K} PF {op.pf_mode} ({op.pf_storage}), {op.device_bytes() / 1e9:.1f} GB, "
        f"set-up {setup:.1f} s")
    if not args.no_autotune:
        tuning = op.autotune(subset=0)
        log(f"[bench] {precision}: autotune " + "; ".join(f"{k}: {v[0]}" for k, v in tuning.items()))
    solver = SPDHGPoisson(op, subsets=M, seed=0)
    solver.graphs = args.graphs
    solver.start(np.zeros(op.coeff_shape, np.float32), b)
    torch.cuda.synchronize()
    solver.advance(args.warmup)
    torch.cuda.synchronize()
    ts = time.perf_counter()
    ms_step, ktimes = timed_steps(solver, args.steps, per_kernel=not args.no_kernel_timing)
    wall = time.perf_counter() - ts
    log(f"[bench] {precision}: timed {args.steps} steps: {ms_step:.3f} ms/step (wall {wall:.3f} s)")
    if ktimes:
        log(f"[bench] {precision}: kernel ms: " + ", ".join(f"{k}={v:.4f}" for k, v in ktimes.items()))
    # continue the same run (untimed) to the step count the precisions' iterates are compared at
    n_done = args.warmup + args.steps
    if args.compare_steps > n_done:
        solver.advance(args.compare_steps - n_done)
        n_done = args.compare_steps
    x_iter = solver.xk.clone()
    solver.finish()
    # full forward / adjoint (API layout, median of 10 after 2)
    xg = torch.from_numpy(xt).cuda()
    bg = torch.from_numpy(np.asarray(b, np.float32)).cuda()
    data = torch.empty(op.data_shape, dtype=torch.float32, device="cuda")
    out = torch.empty(op.coeff_shape, dtype=torch.float32, device="cuda")
    tf, ta = [], []
    for _ in range(0 if args.no_op_timing else 12):
        a0, a1, a2 = (torch.cuda.Event(enable_timing=True) for _ in range(3))
        a0.record()
        op.direct_cl(xg, data)
        a1.record()
        op.adjoint_cl(bg, out)
        a2.record()
        torch.cuda.synchronize()
        tf.append(a0.elapsed_time(a1))
        ta.append(a1.elapsed_time(a2))
    if leg is not None and precision == "fp32" and not args.no_cpu_baseline:
        leg["cpu"], leg["sample"] = cpu_baseline_leg(args, op, b, M)
    fp64 = fp64_errors(op, leg.get("sample")) if leg is not None else None
    res = {"op": op, "M": M, "ms_step": ms_step, "op_nnz": op.sparse_nnz, "counts": b, "fp64": fp64,
           "tuning": {k: v[0] for k, v in (op.tuning or {}).items()},
           "tuning_ms": {k: {c: round(t, 4) for c, t in v[1].items()} for k, v in (op.tuning or {}).items()},
           "ktimes": ktimes, "setup_s": setup,
           "t_fwd": float(np.median(tf[2:])) if tf else None, "t_adj": float(np.median(ta[2:])) if ta else None,
           "fwd_out": data, "adj_out": out, "x_iter": x_iter, "steps_total": n_done}
    return res


def run_c3(args):
    """The multi-slice workload: 64 slices' SPDHG steps batched (BatchedSPDHGPoisson: every kernel of a step
    launched once for all slices), the PF as BF16 hi/lo blocks of the non-zeros (bf16_residual), the steps replayed
    from HIP graphs (every ordered subset pair captured before the timed run), projector variants autotuned on the
    batch. The timed run is `steps` subset steps (each slice's solve: 100 SPDHG iterations at the default). Reports
    slices/s, the p50 / p95 of the batched step, the per-slice solve time (every slice of the batch finishes with the
    batch) and, for reference, the p50 / p95 of the slices' solve times run one at a time (SPDHGPoisson on the same
    operators); the iterates of 3 slices against FP32 single-slice runs (CSR PF) with the same subset order. One
    JSON line."""
    import torch
    from diffractom_amd import SinglePhaseForwardOperator, SPDHGPoisson
    from diffractom_amd.batched import BatchedSPDHGPoisson
    from diffractom_amd.workloads import C3, c3_slices, poisson_counts
    t0 = time.perf_counter()
    cfg, mat, slices = c3_slices(args.slices)
    n_all = len(slices)
    q = "bf16_residual"
    ops, bs, masks, grids = [], [], [], []
    for i, (grid, mask, truth) in enumerate(slices):
        op = SinglePhaseForwardOperator(cfg, mat, grid, max_gb=1024.0, normalized=True, pf_mode="sparse",
                                        pf_precision=q)
        b, _ = poisson_counts(op.direct(truth), C3["mean_counts"], C3["seed"] + i)
        ops.append(op)
        bs.append(b)
        masks.append(mask)
        grids.append(grid)
        if (i + 1) % 16 == 0:
            log(f"[bench] al_c3: {i + 1} slices set up ({time.perf_counter() - t0:.0f} s)")
    Ks = [op.K for op in ops]
    log(f"[bench] al_c3: {n_all} slices set up ({time.perf_counter() - t0:.0f} s, "
        f"{torch.cuda.memory_allocated() / 1e9:.0f} GB, PF {sum(op.device_bytes() for op in ops) / 1e9:.1f} GB)")
    bat = BatchedSPDHGPoisson(ops, subsets=C3["subsets"], support=masks, seed=0, graphs=not args.no_graphs)
    bat.start(None, bs)
    if not args.no_autotune:
        tuning = bat.autotune()
        log("[bench] al_c3: autotune " + "; ".join(f"{k}: {v[0]}" for k, v in tuning.items()))
    torch.cuda.synchronize()
    bat.advance(args.warmup)
    torch.cuda.synchronize()
    bat.finish()
    # the timed solve: a new batch from x = 0 with the subset order of seed 0 (the warm-up drew from the first's)
    del bat
    bat = BatchedSPDHGPoisson(ops, subsets=C3["subsets"], support=masks, seed=0, graphs=not args.no_graphs)
    bat.start(None, bs)
    if bat.graphs:
        log(f"[bench] al_c3: {bat.capture_all()} step graphs captured")
    torch.cuda.synchronize()
    evs = []
    w0, w1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    w0.record()
    for _ in range(args.steps):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        bat.advance(1)
        e1.record()
        evs.append((e0, e1))
    w1.record()
    torch.cuda.synchronize()
    ms = np.array([a.elapsed_time(b) for a, b in evs])
    run_s = w0.elapsed_time(w1) * 1e-3
    log(f"[bench] al_c3: {args.steps} batched steps in {run_s:.3f} s ({ms.mean():.3f} ms per step)")
    xs = bat.finish()
    # FP32 single-slice runs of 3 slices (smallest, median, largest K), the same subset order (seed)
    cmp = {}
    for i in sorted({int(np.argmin(Ks)), int(np.argsort(Ks)[len(Ks) // 2]), int(np.argmax(Ks))}):
        op32 = SinglePhaseForwardOperator(cfg, mat, grids[i], max_gb=1024.0, normalized=True, pf_mode="sparse",
                                          pf_precision="fp32")
        sol = SPDHGPoisson(op32, subsets=C3["subsets"], support=masks[i], seed=0)
        sol.start(torch.zeros(op32.coeff_shape, dtype=torch.float32, device="cuda"), bs[i])
        sol.advance(args.steps)
        x32 = sol.finish()
        xb = xs[i]
        cmp[f"slice {i} (K {Ks[i]})"] = float(torch.linalg.vector_norm((xb - x32).double()) /
                                              torch.linalg.vector_norm(x32.double()))
        del op32, sol, x32
    log("[bench] al_c3: x vs FP32 after the solve: " + ", ".join(f"{k} {v:.2e}" for k, v in cmp.items()))
    del bat, xs
    torch.cuda.empty_cache()
    # the slices one at a time (their own single-slice solver and launches)
    solo = []
    for i, op in enumerate(ops):
        sol = SPDHGPoisson(op, subsets=C3["subsets"], support=masks[i], seed=0)
        sol.start(torch.zeros(op.coeff_shape, dtype=torch.float32, device="cuda"), bs[i])
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        sol.advance(args.steps)
        e1.record()
        torch.cuda.synchronize()
        solo.append(e0.elapsed_time(e1) * 1e-3)
        sol.finish()
        del sol
    solo = np.array(solo)
    log(f"[bench] al_c3: one slice at a time: {solo.sum():.2f} s for all, p50 {np.percentile(solo, 50):.3f} s, "
        f"p95 {np.percentile(solo, 95):.3f} s per slice")
    line = {"metric": "SPDHG multi-slice batch: slices/s", "value": round(n_all / run_s, 4),
            "unit": f"slices/s ({args.steps} subset steps per slice)", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": round(float(ms.mean()), 4), "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": DTYPE[q],
            "data": "synthetic (jittered-grain grids, Voronoi grains in disc masks, Poisson counts)",
            "config": {"workload": f"al_c3: {n_all} slices in one batch, 128x128 px, 180 omega x 131 x 180 eta x 5 "
                                   f"rings, K {min(Ks)}..{max(Ks)}, disc masks r 15..60 px, M=10, {PF_DESC[q]}, "
                                   f"{'HIP graphs' if not args.no_graphs else 'eager launches'}",
                       "slices": n_all, "K_sum": int(sum(Ks)), "precision": q},
            "step_ms_p50": round(float(np.percentile(ms, 50)), 4), "step_ms_p95": round(float(np.percentile(ms, 95)), 4),
            "slice_solve_s_batched": round(run_s, 4),
            "slice_solve_s_alone_p50": round(float(np.percentile(solo, 50)), 4),
            "slice_solve_s_alone_p95": round(float(np.percentile(solo, 95)), 4),
            "slices_per_s_alone": round(n_all / float(solo.sum()), 4),
            "rel_l2_x_vs_fp32": cmp}
    print(json.dumps(line), flush=True)


def run_local(args):
    """The local-basis workload (al_local: the al_c1 scan, grid and truth; every voxel allows the orientations within
    3.5 deg of its grain's, grown by 2 voxels): SPDHG subset steps per second of the local operator
    (LocalSinglePhaseForwardOperator, --precision PF products) with per-kernel times of the timed steps, against the
    global operator on the same orientations (full images); forward / adjoint of the local operator against the
    global one and against the FP64 oracle on 4 angle rows. One JSON line."""
    import torch
    from diffractom_amd import MatrixGrid, SinglePhaseForwardOperator, SPDHGPoisson
    from diffractom_amd.workloads import AL_RINGS, al_local_problem
    from oracle.local import OracleLocalOperator
    from oracle.model import OracleOperator
    q = args.precision
    t0 = time.perf_counter()
    op, xt, b, p = al_local_problem(precision=q)
    setup_s = time.perf_counter() - t0
    sup = op.support
    log(f"[bench] al_local: {sup.n_pairs} pairs, {op.K} of {op.K_basis} orientations, {op.n_patches} patches, "
        f"{op.n_coef} coefficients, strips {100 * op.strip_fraction:.1f} % of the sinograms; set up {setup_s:.1f} s")
    M = p["subsets"]
    if not args.no_autotune:
        op.set_omega_subsets(M)
        tuning = op.autotune()
        log("[bench] al_local: autotune " + "; ".join(f"{k}: {v[0]}" for k, v in tuning.items()))
    dev = torch.device("cuda")
    bd = torch.from_numpy(b).to(dev)

    def solve(o, x0):
        sol = SPDHGPoisson(o, subsets=M, seed=0)
        sol.graphs = args.graphs
        sol.start(x0, bd)
        sol.advance(args.warmup)
        torch.cuda.synchronize()
        ms, _ = timed_steps(sol, args.steps, per_kernel=False)
        _, kms = timed_steps(sol, min(args.steps, 20), per_kernel=not args.no_kernel_timing)
        sol.finish()
        return ms, kms

    ms_l, kms_l = solve(op, torch.zeros(op.coeff_shape, dtype=torch.float32, device=dev))
    log(f"[bench] al_local: local {ms_l:.4f} ms/step; kernel ms: " +
        ", ".join(f"{k} {v:.4f}" for k, v in (kms_l or {}).items()))
    if args.no_compare:
        return
    # the global operator on the same orientations (operator order), full images
    U = op.grid_rot_cpu
    sig = float(np.asarray(op.sigma_cpu)[0])
    gop = SinglePhaseForwardOperator(op.cfg, op.material, MatrixGrid(U, sig), max_gb=1024.0, normalized=True,
                                     pf_mode="sparse", pf_precision=q)
    if not args.no_autotune:
        gop.set_omega_subsets(M)
        gop.autotune()
    ms_g, kms_g = solve(gop, torch.zeros(gop.coeff_shape, dtype=torch.float32, device=dev))
    log(f"[bench] al_local: global (same {gop.K} orientations) {ms_g:.4f} ms/step")
    # parity: the local operator against the global one on the same coefficients, and against the FP64 oracle
    inv = np.full(op.K_basis, -1, np.int64)
    inv[op.orientation_index] = np.arange(op.K)
    k, iy, ix, pi = op._pairs()
    xd = np.zeros(gop.coeff_shape, np.float32)
    xd[inv[k], iy, ix] = xt[pi]
    rng = np.random.default_rng(1)
    y = rng.random(op.data_shape).astype(np.float32)
    d_l, d_g = op.direct(xt), gop.direct(xd)
    a_l, a_g = op.adjoint(y), gop.adjoint(y)
    rel = lambda u, v: float(np.linalg.norm((np.asarray(u, np.float64) - v).ravel()) /
                             max(np.linalg.norm(np.asarray(v, np.float64).ravel()), 1e-300))
    vs_global = {"forward": rel(d_l, d_g), "adjoint": rel(a_l, _gather(a_g, inv[k], iy, ix, pi, op.n_coef))}
    rows = [0, 45, 90, 135]
    t1 = time.perf_counter()
    base = OracleOperator(op.cfg, (4.05, 4.05, 4.05, 90, 90, 90), "cubic", AL_RINGS, 0.2, U, sig,
                          orthonormalize=False)
    orc = OracleLocalOperator(base, inv[k], iy, ix, pi, op.n_coef)
    base.set_omega_subsets([np.asarray(rows)])
    ym = np.zeros_like(y)
    ym[rows] = y[rows]
    vs_oracle = {"forward": rel(d_l[rows], orc.direct_rows(xt, rows)),
                 "adjoint": rel(op.adjoint(ym), orc.adjoint(y, subset=0)),
                 "sample": f"angle rows {rows}, all {op.K} orientations"}
    oracle_s = time.perf_counter() - t1
    log(f"[bench] al_local: vs global {vs_global}, vs FP64 oracle {vs_oracle} ({oracle_s:.1f} s)")
    line = {"metric": "SPDHG-Poisson subset steps/s on a local basis (al_c1 scan)", "value": round(1e3 / ms_l, 3),
            "unit": "SPDHG subset steps/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": round(ms_l, 4), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": DTYPE[q], "data": "synthetic (al_c1 Voronoi grains, Poisson counts at a mean of 1000)",
            "config": {"workload": f"al_local: al_c1 scan and 3-degree grid (K {op.K_basis}), support = "
                                   f"orientations within {p['tol_deg']} deg of each voxel's grain, grown by "
                                   f"{p['dilate']} voxels: {sup.n_pairs} pairs, {op.K} orientations, "
                                   f"{op.n_patches} patches, {op.n_coef} coefficients; M={M}; {PF_DESC[q]}",
                       "K_used": op.K, "n_coef": op.n_coef, "precision": q},
            "setup_s": round(setup_s, 2), "kernel_ms": kms_l,
            "global_same_orientations": {"ms_per_step": round(ms_g, 4), "steps_per_s": round(1e3 / ms_g, 3),
                                         "kernel_ms": kms_g, "coefficients": int(np.prod(gop.coeff_shape))},
            "speedup_vs_global": round(ms_g / ms_l, 3),
            "rel_l2_vs_global": vs_global, "rel_l2_vs_fp64_oracle": vs_oracle}
    print(json.dumps(line), flush=True)


def run_extras(args):
    """The widened paths on the al_c1 workload (K = 29713, 99 x 99 pixels, 180 angles; the al_c1 truth and counts),
    one JSON line of times (HIP events, the median of 3 calls after one warm-up call):
      parameter gradients: PFGradient (rotation, width, eta offset, axis tilt, Bragg angles; with the Gauss-Newton
        blocks) and ProjGradient (per-angle shifts, centre of rotation) of <r, A c>, r = 1 - b / (A c + 1e-3);
      Fourier-slice projector (projector="fft"): full forward and adjoint against the native projector's, and the
        relative difference of the two forwards (the Fourier model differs from the native footprint by design);
      host streaming (HostStreamer): forward and adjoint of (K, Ny, Nx) host arrays against the device products."""
    import torch
    from diffractom_amd import SinglePhaseForwardOperator
    from diffractom_amd.gradients import PFGradient, ProjGradient
    from diffractom_amd.streaming import HostStreamer
    from diffractom_amd.workloads import make_grid
    q = args.precision
    op, b, xt, p = build_problem("al_c1", precision=q)
    dev = torch.device("cuda")
    c = torch.from_numpy(xt).to(dev)
    bd = torch.from_numpy(b).to(dev)
    ax = op.direct(c)
    r = (1.0 - bd / (ax + 1e-3)).contiguous()

    def med(fn, n=3, label=""):
        t0 = time.perf_counter()
        fn()
        torch.cuda.synchronize()
        log(f"[bench] extras: {label} warm-up call {time.perf_counter() - t0:.2f} s")
        ts = []
        for _ in range(n):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            fn()
            e1.record()
            torch.cuda.synchronize()
            ts.append(e0.elapsed_time(e1))
        return float(np.median(ts))

    out = {}
    pg, pgh, jg = PFGradient(op), PFGradient(op, hessian=True), ProjGradient(op)
    out["pf_gradient_ms"] = round(med(lambda: pg(c, r), label="pf_gradient"), 3)
    out["pf_gradient_gauss_newton_ms"] = round(med(lambda: pgh(c, r), label="pf_gradient_gauss_newton"), 3)
    out["proj_gradient_ms"] = round(med(lambda: jg(c, r), label="proj_gradient"), 3)
    log(f"[bench] extras: PFGradient {out['pf_gradient_ms']} ms, with Gauss-Newton blocks "
        f"{out['pf_gradient_gauss_newton_ms']} ms, ProjGradient {out['proj_gradient_ms']} ms")
    data = torch.empty(op.data_shape, dtype=torch.float32, device=dev)
    img = torch.empty(op.coeff_shape, dtype=torch.float32, device=dev)
    out["native_forward_ms"] = round(med(lambda: op.direct_cl(c, data), label="native_forward"), 3)
    out["native_adjoint_ms"] = round(med(lambda: op.adjoint_cl(bd, img), label="native_adjoint"), 3)
    # host streaming of (K, Ny, Nx) host arrays
    hs = HostStreamer(op)
    xh = np.ascontiguousarray(xt, dtype=np.float32)
    oh = np.empty_like(xh)
    gb = xh.nbytes / 1e9
    out["stream_forward_ms"] = round(med(lambda: hs.forward(xh, data), label="stream_forward"), 3)
    out["stream_adjoint_ms"] = round(med(lambda: hs.adjoint(bd, oh), label="stream_adjoint"), 3)
    out["stream_host_GBps"] = {"forward": round(gb / out["stream_forward_ms"] * 1e3, 1),
                               "adjoint": round(gb / out["stream_adjoint_ms"] * 1e3, 1)}
    hs.release()
    log(f"[bench] extras: native forward {out['native_forward_ms']} / adjoint {out['native_adjoint_ms']} ms; "
        f"host-streamed {out['stream_forward_ms']} / {out['stream_adjoint_ms']} ms ({gb:.2f} GB host arrays)")
    d_native = data.clone()
    del op, pg, pgh, jg
    torch.cuda.empty_cache()
    # the Fourier-slice projector on the same geometry and PF
    grid, mat, cfg = make_grid("al_c1")
    t0 = time.perf_counter()
    fop = SinglePhaseForwardOperator(cfg, mat, grid, max_gb=1024.0, normalized=True, pf_mode="sparse",
                                     pf_precision=q, projector="fft")
    out["fft_setup_s"] = round(time.perf_counter() - t0, 1)
    out["fft_forward_ms"] = round(med(lambda: fop.direct_cl(c, data), label="fft_forward"), 3)
    out["fft_adjoint_ms"] = round(med(lambda: fop.adjoint_cl(bd, img), label="fft_adjoint"), 3)
    fop.direct_cl(c, data)
    out["fft_vs_native_forward_rel_l2"] = float(torch.linalg.vector_norm((data - d_native).double()) /
                                                torch.linalg.vector_norm(d_native.double()))
    log(f"[bench] extras: fft forward {out['fft_forward_ms']} / adjoint {out['fft_adjoint_ms']} ms, "
        f"vs native forward {out['fft_vs_native_forward_rel_l2']:.2e}")
    line = {"metric": "widened paths on al_c1: parameter gradients, Fourier-slice projector, host streaming",
            "value": out["pf_gradient_ms"], "unit": "ms per PFGradient call", "n_gpus": 1, "steps": 3, "warmup": 1,
            "ms_per_step": out["pf_gradient_ms"], "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
            "dtype": DTYPE[q], "data": "synthetic (al_c1 Voronoi grains, Poisson counts at a mean of 1000)",
            "config": {"workload": f"al_c1: K={xt.shape[0]}, 99x99 px, 180 omega x 103 x 180 eta x 5 rings; "
                                   f"{PF_DESC[q]}", "precision": q},
            **out}
    print(json.dumps(line), flush=True)


def _gather(c, k, iy, ix, pi, n):
    """(n,) flat local coefficients of a (K, Ny, Nx) array at the pairs (zero elsewhere)."""
    out = np.zeros(n, np.float32)
    out[pi] = c[k, iy, ix]
    return out


def run_c2(args):
    """The large hexagonal workload (Ti, 256 x 256 px, 360 omega, K = 59,440): the PF generated per call in fp16
    blocks on the matrix cores (--c2-precision fp16, the configuration's own form; bf16_residual possible) or stored
    as an FP32 CSR of about 4 G non-zeros (--c2-precision fp32). Full forward / adjoint times and GB/s, SPDHG subset
    steps per second with per-kernel times of the timed steps, the PF products' MFMA figures, and the forward rows /
    block adjoint against the FP64 oracle on a sample (2 angles x 256 orientations). One JSON line."""
    import torch
    from diffractom_amd import SPDHGPoisson
    from diffractom_amd.workloads import TI_RINGS, c2_problem
    q = args.c2_precision
    t0 = time.perf_counter()
    op, xk, b, p = c2_problem(pf_mode="sparse" if q == "fp32" else "generated", pf_precision=q)
    torch.cuda.synchronize()
    setup = time.perf_counter() - t0
    log(f"[bench] ti_c2 {q}: K={op.K} nnz={op.sparse_nnz} fill={op.sparse_fill:.4f} batches={len(op.batches)}, "
        f"PF {op.pf_mode} ({op.pf_storage}), {op.device_bytes() / 1e9:.1f} GB, set-up {setup:.1f} s")
    npix = op.Nx * op.Ny
    M = p["subsets"]
    op.set_omega_subsets(M)
    if not args.no_autotune:
        tuning = op.autotune(subset=0, reps=3)
        log("[bench] ti_c2: autotune " + "; ".join(f"{k}: {v[0]}" for k, v in tuning.items()))
    data = torch.empty(op.data_shape, dtype=torch.float32, device="cuda")
    out = torch.empty_like(xk).view(-1)
    tf, ta = [], []
    for _ in range(3):
        a0, a1, a2 = (torch.cuda.Event(enable_timing=True) for _ in range(3))
        a0.record()
        op.direct_k(xk.view(-1), data)
        a1.record()
        op.adjoint_k(data, out)
        a2.record()
        torch.cuda.synchronize()
        tf.append(a0.elapsed_time(a1))
        ta.append(a1.elapsed_time(a2))
    del out, data
    # algorithmic bytes of a full product: the image and the sinogram once, plus the stored CSR's 6 B per non-zero
    # (4 B value + 2 B column); the generated blocks are made and used within the call (written and read back once
    # from the scratch, 4 B / 6 B per non-zero, reported as entry traffic, not as algorithmic bytes)
    fbytes = 4.0 * op.K * npix + 4.0 * op.N_Omega * op.My * op.N_seg + (6.0 * op.sparse_nnz if q == "fp32" else 0.0)
    log(f"[bench] ti_c2: forward {min(tf):.1f} ms, adjoint {min(ta):.1f} ms")
    del xk
    torch.cuda.empty_cache()
    fp64 = c2_oracle_errors(op, b) if not args.no_cpu_baseline else None
    solver = SPDHGPoisson(op, subsets=M, seed=0)
    x0 = torch.zeros(op.coeff_shape, dtype=torch.float32, device="cuda")
    solver.start(x0, b)
    torch.cuda.synchronize()
    solver.advance(args.warmup)
    torch.cuda.synchronize()
    ms, kt = timed_steps(solver, args.steps, per_kernel=not args.no_kernel_timing)
    log(f"[bench] ti_c2: {ms:.2f} ms/step; kernel ms: " + ", ".join(f"{k}={v:.3f}" for k, v in (kt or {}).items()))
    mf = mfma_figures(op, M, kt) if kt and op.gen_batches is None else None
    if kt and op.gen_batches is not None:
        # the generated blocks: MFMA issued per product (3 passes with the residual, 1 otherwise)
        passes = 3 if q.endswith("residual") else 1
        eb = 4.0 if q == "fp16" else 6.0             # 2 B position + 2 B fp16 value, or a bf16 hi/lo pair
        Rm = op.N_Omega / M
        issued = 2.0 * Rm * (-(-op.My // 16) * 16) * (-(-op.N_seg // 32) * 32) * \
            sum(-(-bt["K_batch"] // 32) * 32 for bt in op.batches) * passes
        mf = {k: {"mfma_TFLOPs_issued": round(issued / (kt[k] * 1e-3) / 1e12, 2),
                  "mfma_frac_issued": round(issued / (kt[k] * 1e-3) / 1e12 / MFMA_BF16_PEAK_TFLOPS, 4),
                  "entry_bytes": eb * op.sparse_nnz / M,
                  "entry_GBps": round(eb * op.sparse_nnz / M / (kt[k] * 1e-3) / 1e9, 1)}
              for k in ("pf_forward", "pf_adjoint") if k in kt}
    desc = "FP32 CSR PF" if q == "fp32" else f"PF generated per call, {q} blocks on the matrix cores"
    line = {"metric": f"SPDHG-Poisson iters/s (subset steps/s), hexagonal Ti, {desc}",
            "value": round(1000.0 / ms, 4), "unit": "SPDHG subset steps/s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": round(ms, 3), "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": DTYPE[q], "data": "synthetic (300 Voronoi grains + 10 % fibre background)",
            "config": {"workload": f"ti_c2: Ti 8 rings, 360 omega x 261 x 360 eta, 256x256 px, K={op.K}, sigma 2.5 "
                                   f"deg, M={M}, {desc} (nnz {op.sparse_nnz})",
                       "K": op.K, "nnz": op.sparse_nnz, "precision": q, "pf_mode": op.pf_mode},
            "setup_s": round(setup, 1), "device_GB": round(op.device_bytes() / 1e9, 2),
            "forward_ms": round(min(tf), 2), "adjoint_ms": round(min(ta), 2),
            "forward_GBps": round(fbytes / (min(tf) * 1e-3) / 1e9, 1),
            "adjoint_GBps": round(fbytes / (min(ta) * 1e-3) / 1e9, 1),
            "kernel_ms": {k: round(v, 4) for k, v in (kt or {}).items()}, "matrix_cores": mf,
            "rel_err_vs_fp64_oracle": fp64}
    print(json.dumps(line), flush=True)


def c2_oracle_errors(op, b):
    """Relative L2 errors of the C2 operator's forward rows (angles 10 and 100) for coefficients on 256 orientations
    straddling the first two orientation batches, and of the block adjoint of the counts on those rows, against the
    FP64 oracle (oracle/model.py)."""
    import torch
    try:
        from diffractom_amd.workloads import TI_RINGS
        from oracle.model import OracleOperator
        kb0 = op.batches[0]["K_batch"]
        ks = np.arange(max(kb0 - 128, 0), min(kb0 + 128, op.K))
        rows = np.array([10, 100])
        orc = OracleOperator(op.cfg, (2.95, 2.95, 4.68, 90, 90, 120), "hexagonal", TI_RINGS, 0.2,
                             op.grid_rot_cpu[ks], float(op.sigma_cpu[0]), orthonormalize=False)
        xb = np.random.default_rng(3).random((len(ks), op.Ny, op.Nx)) * op.support_mask()[None]
        ref_f = orc.direct_rows(xb, rows)
        orc.set_omega_subsets([rows])
        yb = np.zeros(op.data_shape)
        rt = torch.from_numpy(rows).cuda()
        yb[rows] = b[rt].double().cpu().numpy()
        ref_a = orc.adjoint(yb, subset=0)
        x = torch.zeros(op.coeff_shape, dtype=torch.float32, device="cuda")
        x[int(ks[0]):int(ks[-1]) + 1] = torch.from_numpy(xb.astype(np.float32)).cuda()
        fwd = op.direct(x)[rt].double().cpu().numpy()
        del x
        y = torch.zeros(op.data_shape, dtype=torch.float32, device="cuda")
        y[rt] = b[rt]
        adj = op.adjoint(y)[int(ks[0]):int(ks[-1]) + 1].double().cpu().numpy()
        del y
        torch.cuda.empty_cache()
        e = lambda a, r: float(np.linalg.norm(a - r) / np.linalg.norm(r))
        out = {"forward": e(fwd, ref_f), "adjoint": e(adj, ref_a),
               "sample": f"angle rows {rows.tolist()} x orientations {int(ks[0])}..{int(ks[-1])}"}
        log(f"[bench] ti_c2: FP64 oracle sample: forward {out['forward']:.2e}, adjoint {out['adjoint']:.2e}")
        return out
    except Exception as exc:
        log(f"[bench] ti_c2: FP64 oracle sample failed: {exc!r}")
        return None


def run_osem(args):
    """OSEMPoisson on al_c0 / al_c1 (the same counts and operator as the SPDHG line, --precision's PF product):
    warm-up, then `steps` timed subset iterations (forward of the subset, EM ratio, subset PF adjoint, the
    backprojection with the multiplicative update) with x and the M sensitivities resident in HBM in the
    orientation-fastest layout. One JSON line: subset iterations/s, and the deviance before and after."""
    import torch
    from diffractom_amd import OSEMPoisson
    op, b, _, p = build_problem(args.config, precision=args.precision)
    M = p["subsets"]
    if not args.no_autotune:
        tuning = op.autotune(subset=0)
        log(f"[bench] osem {args.precision}: autotune " + "; ".join(f"{k}: {v[0]}" for k, v in tuning.items()))
    em = OSEMPoisson(op, subsets=M, seed=0)
    em.start(np.zeros(op.coeff_shape, np.float32), b)
    f0 = em.deviance()
    order = []
    while len(order) < args.warmup + args.steps:
        order += [int(m) for m in em.rng.permutation(M)]
    for m in order[: args.warmup]:
        em.iterate(m)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for m in order[args.warmup: args.warmup + args.steps]:
        em.iterate(m)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / args.steps
    f1 = em.deviance()
    log(f"[bench] osem {args.precision}: {ms:.3f} ms per subset iteration (EM update fused into the backprojection: "
        f"{op.em_fused(0)}); deviance {f0:.6g} -> {f1:.6g}")
    line = {"metric": "OSEM-Poisson subset iterations/s", "value": round(1000.0 / ms, 3),
            "unit": "OSEM subset iterations/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": round(ms, 4), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": DTYPE[args.precision], "data": "synthetic (Voronoi grains, Poisson counts at a mean of 1000)",
            "config": {"workload": f"{args.config}: K={op.K}, M={M} subsets, PF product {PF_DESC[args.precision]}",
                       "K": op.K, "subsets": M, "precision": args.precision},
            "deviance": {"start": f0, "after": f1, "iterations": args.warmup + args.steps}}
    print(json.dumps(line), flush=True)


def run_pdhg(args):
    """PDHGPoisson on al_c0 / al_c1 (the SPDHG line's counts and operator, --precision's PF product): warm-up, then
    `steps` timed iterations over all angles (dual update, PF adjoint, the backprojection with the primal step and
    the extrapolation, forward of x_bar, the A x average, the float64 deviance) with x, x_bar and T resident in HBM.
    One JSON line: iterations/s and the deviance before and after."""
    import torch
    from diffractom_amd import PDHGPoisson
    op, b, _, p = build_problem(args.config, precision=args.precision)
    if not args.no_autotune:
        tuning = op.autotune(subset=0)
        log(f"[bench] pdhg {args.precision}: autotune " + "; ".join(f"{k}: {v[0]}" for k, v in tuning.items()))
    pd = PDHGPoisson(op)
    pd.start(np.zeros(op.coeff_shape, np.float32), b)
    f0 = pd._deviance(pd.Ax)
    for _ in range(args.warmup):
        pd.iterate()
        pd._deviance(pd.Ax)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(args.steps):
        pd.iterate()
        f1 = pd._deviance(pd.Ax)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / args.steps
    log(f"[bench] pdhg {args.precision}: {ms:.3f} ms per iteration (primal step fused into the backprojection: "
        f"{op.em_fused(None)}); deviance {f0:.6g} -> {f1:.6g}")
    line = {"metric": "PDHG-Poisson iterations/s", "value": round(1000.0 / ms, 3), "unit": "PDHG iterations/s",
            "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms, 4),
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": DTYPE[args.precision],
            "data": "synthetic (Voronoi grains, Poisson counts at a mean of 1000)",
            "config": {"workload": f"{args.config}: K={op.K}, all {op.N_Omega} angles per iteration, PF product "
                                   f"{PF_DESC[args.precision]}", "K": op.K, "precision": args.precision},
            "deviance": {"start": f0, "after": f1, "iterations": args.warmup + args.steps}}
    print(json.dumps(line), flush=True)


def run_fista(args):
    """FISTAPoisson on al_c0 / al_c1 (the SPDHG line's counts and operator, --precision's PF product): one run of
    warmup + steps iterations (the weighted-least-squares start phase and the metric refreshes of the schedule
    included); the timed figure is the wall clock of its last `steps` iterations (from iter_stats: every iteration
    waits for its descent test on the host, as the reference's). One JSON line: iterations/s and the objective."""
    import torch
    from diffractom_amd import FISTAPoisson
    op, b, _, p = build_problem(args.config, precision=args.precision)
    if not args.no_autotune:
        tuning = op.autotune(subset=0)
        log(f"[bench] fista {args.precision}: autotune " + "; ".join(f"{k}: {v[0]}" for k, v in tuning.items()))
    fi = FISTAPoisson(op)
    x = torch.zeros(op.coeff_shape, dtype=torch.float32, device="cuda")
    fi.run(x, b, args.warmup + args.steps)
    torch.cuda.synchronize()
    st = fi.iter_stats
    ms = (st[-1]["time"] - st[args.warmup - 1]["time"]) * 1000.0 / args.steps
    # the objective from the first iteration of the Poisson phase on (the start phase minimises the weighted least
    # squares approximation, another functional)
    pois = [s for s in st if s["phase"] == "poisson"]
    f0 = pois[0]["obj"] if pois else float("nan")
    it0 = pois[0]["iter"] if pois else None
    bt = sum(s["backtracks"] for s in st[args.warmup:])
    refreshes = sum(bool(s["metric_refresh"]) for s in st[args.warmup:])
    log(f"[bench] fista {args.precision}: {ms:.3f} ms per iteration (wall clock over iterations "
        f"{args.warmup + 1}..{args.warmup + args.steps}: {bt} backtracks, {refreshes} metric refreshes); Poisson "
        f"objective at iteration {it0}: {f0:.6g} -> {fi.final_stats['final_obj']:.6g}")
    line = {"metric": "FISTA-Poisson iterations/s", "value": round(1000.0 / ms, 3), "unit": "FISTA iterations/s",
            "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms, 4),
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": DTYPE[args.precision],
            "data": "synthetic (Voronoi grains, Poisson counts at a mean of 1000)",
            "config": {"workload": f"{args.config}: K={op.K}, all {op.N_Omega} angles per iteration, PF product "
                                   f"{PF_DESC[args.precision]}", "K": op.K, "precision": args.precision},
            "objective": {"poisson_phase_start": f0, "poisson_phase_start_iter": it0,
                          "after": fi.final_stats["final_obj"], "backtracks": bt, "metric_refreshes": refreshes}}
    print(json.dumps(line), flush=True)


def run_c4(args):
    """Differentiation at the C4 size: 8 slices of 128 x 128 px (180 omega x 131 x 180 eta x 5 rings, K = 5000,
    sigma 2 deg, BF16 residual PF), loss = Poisson NLL(f_a * A c; b) + 1e-3 TV(c), gradients with respect to c and
    the per-angle scales f_a ~ U(0.9, 1.1) through the autograd Function (backward = the HIP adjoint). One JSON line:
    loss + gradient evaluations per second (all 8 slices)."""
    import torch
    from diffractom_amd import SinglePhaseForwardOperator
    from diffractom_amd.differentiable import IntensityLayer, forward_project, poisson_nll, total_variation
    from diffractom_amd.workloads import C3, jittered_grid, al_material, scan_cfg, poisson_counts, voronoi_truth
    from diffractom_amd.grid import MatrixGrid
    cfg = scan_cfg(128, 131, 180, 180)
    rng = np.random.default_rng(4)
    slices = []
    for i in range(8):
        R = jittered_grid(5000, 100, 1.0, "cubic", int(rng.integers(1 << 31)))
        op = SinglePhaseForwardOperator(cfg, al_material(), MatrixGrid(R, np.deg2rad(2.0)), max_gb=1024.0,
                                        normalized=True, pf_mode="dense", pf_precision="bf16_residual")
        if not args.no_autotune:
            op.autotune(subset=0, reps=3)
        xt = voronoi_truth(op.K, 128, 60, 55.0, int(rng.integers(1 << 31)), coeff=(0.5, 1.5),
                           orientations=rng.choice(100, 60, replace=False))
        b, _ = poisson_counts(op.direct(xt), 1000.0, i)
        lay = IntensityLayer(op.data_shape, op.N_peaks)
        prm = lay.params0(device="cuda")
        prm["angle_scale"] = torch.from_numpy(rng.uniform(0.9, 1.1, op.N_Omega)).cuda().requires_grad_(True)
        c = torch.full(op.coeff_shape, 0.01, dtype=torch.float32, device="cuda").requires_grad_(True)
        slices.append((op, lay, prm, c, torch.from_numpy(b).cuda()))
    log("[bench] al_c4: 8 slices set up")

    def evaluate():
        total = 0.0
        for op, lay, prm, c, bt in slices:
            c.grad = None
            prm["angle_scale"].grad = None
            loss = poisson_nll(lay.forward(forward_project(c, op), prm), bt) + 1e-3 * total_variation(c)
            loss.backward()
            total = total + loss.detach()
        return total

    for _ in range(args.warmup):
        evaluate()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(args.steps):
        evaluate()
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / args.steps
    line = {"metric": "loss + gradient evaluations/s (8 slices: Poisson NLL + 1e-3 TV, d/dc and d/df_a)",
            "value": round(1000.0 / ms, 3), "unit": "evaluations/s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": round(ms, 3), "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "bf16", "data": "synthetic (jittered-grain grids, Voronoi grains)",
            "config": {"workload": "al_c4: 8 slices, 128x128 px, 180 omega x 131 x 180 eta x 5 rings, K=5000, "
                                   "BF16 residual PF, autograd through the HIP adjoint"}}
    print(json.dumps(line), flush=True)


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this benchmark measures one GPU; sharding the orientations over several GPUs "
                                   "is a design, not a measured path",
                          "n_gpus": args.gpus}), flush=True)
        sys.exit(2)
    import torch
    if args.config in ("al_c3", "ti_c2", "al_c4", "al_local", "al_extras"):
        {"al_c3": run_c3, "ti_c2": run_c2, "al_c4": run_c4, "al_local": run_local,
         "al_extras": run_extras}[args.config](args)
        return
    if args.solver in ("osem", "pdhg", "fista"):
        {"osem": run_osem, "pdhg": run_pdhg, "fista": run_fista}[args.solver](args)
        return
    # fp32 first: its counts are the problem every precision solves
    precisions = [args.precision] if args.only else ["fp32"] + [q for q in ("bf16_residual", "fp16", "bf16")]
    results = {}
    counts = None
    leg = {}
    for q in precisions:
        r = run_path(args, q, counts, leg)
        counts = r["counts"]
        results[q] = r
        if q != args.precision:     # keep the headline path's operator (roofline) only
            r["op"].free_memory()
            r["op"] = None
        torch.cuda.empty_cache()
    head = results[args.precision]
    op = head["op"]
    M = head["M"]
    npix = op.Nx * op.Ny
    K, R, My, CP = op.K, op.N_Omega, op.My, op.N_seg
    paths = {}
    ref = results.get("fp32")
    for q, r in results.items():
        e = {"steps_per_s": round(1000.0 / r["ms_step"], 3), "ms_per_step": round(r["ms_step"], 4),
             "forward_ms": r["t_fwd"] and round(r["t_fwd"], 4), "adjoint_ms": r["t_adj"] and round(r["t_adj"], 4),
             "kernel_ms": {k: round(v, 5) for k, v in (r["ktimes"] or {}).items()}, "tuned": r["tuning"],
             "autotune_candidates_ms": r.get("tuning_ms"),
             "rel_l2_vs_fp64_oracle": r["fp64"]}
        if ref is not None and q != "fp32":
            e["rel_l2_vs_fp32"] = {"forward": rel_l2(r["fwd_out"], ref["fwd_out"]),
                                   "adjoint": rel_l2(r["adj_out"], ref["adj_out"]),
                                   f"x_after_{r['steps_total']}_steps": rel_l2(r["x_iter"], ref["x_iter"])}
        paths[q] = e
    roofline = roofline_all = None
    if head["ktimes"]:
        kb = kernel_bytes(op, M)
        kt = head["ktimes"]
        dom = max(kt, key=kt.get)
        ach = kb[dom] / (kt[dom] * 1e-3) / 1e9
        tuned = head["tuning"] or {}
        roofline = {"kernel": dom, "variant": tuned.get(dom), "bound": "hbm", "achieved": round(ach, 1),
                    "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4),
                    "traffic": measured_traffic(dom, args.config, args.precision, tuned.get(dom)),
                    "avg_ms": round(kt[dom], 5), "algorithmic_bytes": kb[dom]}
        # every step kernel against the same HBM roofline (the forward projection is bound by its image gathers
        # through L2, not by HBM: its algorithmic bytes are the image read once and the sinogram written)
        roofline_all = {k: {"avg_ms": round(kt[k], 5), "algorithmic_bytes": kb[k],
                            "achieved_GBps": round(kb[k] / (kt[k] * 1e-3) / 1e9, 1),
                            "frac": round(kb[k] / (kt[k] * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                            "traffic": measured_traffic(k, args.config, args.precision, tuned.get(k))}
                        for k in kt if k in kb}
        for k, v in (mfma_figures(op, M, kt) or {}).items():
            roofline_all[k]["matrix_cores"] = v

    cpu = leg.get("cpu")
    smp = leg.get("sample")

    ms_step = head["ms_step"]
    value = 1000.0 / ms_step
    # forward / adjoint GB/s: x + data + one CSR of the PF matrix (its non-zeros, counted when the PF was built)
    nnz_all = pf_nnz(op)
    fbytes = 4.0 * K * npix + 4.0 * R * My * CP + 6.0 * nnz_all
    line = {
        "metric": "SPDHG-Poisson iters/s + fwd/adj GB/s, BF16-MFMA vs FP32, <=1e-3 rel err",
        "value": round(value, 3), "unit": "SPDHG subset steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms_step, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": DTYPE[args.precision],
        "rel_err_vs_fp32": paths[args.precision].get("rel_l2_vs_fp32"),
        "rel_err_vs_fp64_oracle": head["fp64"],
        "fp64_oracle": smp and {"seconds": round(smp["seconds"], 2), "kind": "FP64 NumPy/SciPy restatement",
                                "sample": f"forward rows of angles {list(map(int, smp['rows']))} for coefficients "
                                          f"on orientations {int(smp['ks'][0])}..{int(smp['ks'][-1])}, and the "
                                          f"adjoint of the counts on those rows over the same orientations"},
        "pf_nnz": nnz_all,
        "data": "synthetic (Voronoi grains, Poisson counts at a mean of 1000)",
        "config": {"workload": f"{args.config}: Al 5 rings, 180 omega x 103 x 180 eta, 99 x 99 px, K={op.K}, "
                               f"sigma 2 deg, M={M} subsets; PF product {PF_DESC[args.precision]}",
                   "K": op.K, "subsets": M, "precision": args.precision},
        "roofline": roofline,
        "roofline_step_kernels": roofline_all,
        "cpu_baseline": cpu,
        "forward_GBps": head["t_fwd"] and round(fbytes / (head["t_fwd"] * 1e-3) / 1e9, 1),
        "adjoint_GBps": head["t_adj"] and round(fbytes / (head["t_adj"] * 1e-3) / 1e9, 1),
        "paths": paths,
    }
    print(json.dumps(line), flush=True)
    if args.copy_calibration:
        src = torch.empty(1 << 28, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        src.fill_(1.0)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        log("[bench] copy calibration: 1 GiB read, 1 GiB written")


def measured_traffic(kernel, config, precision, variant=None):
    """HBM bytes per launch of `kernel` on this workload from the committed counter summary
    (profiles/traffic_<workload>.json: FETCH_SIZE + WRITE_SIZE per dispatch, scaled by the copy calibration), or
    None (also when the summary was measured on another kernel variant than the one timed here)."""
    path = os.path.join(ROOT, "profiles", f"traffic_{config}.json")
    try:
        with open(path) as f:
            t = json.load(f)
    except OSError:
        return None
    e = t.get("kernels", {}).get(f"{precision}:{kernel}")
    if e is None or (variant is not None and e.get("variant") not in (None, variant)):
        return None
    return float(e["bytes_per_launch"])


DTYPE = {"bf16": "bf16", "bf16_residual": "bf16", "fp16": "f16", "fp32": "f32"}
PF_DESC = {"bf16": "block-sparse BF16 MFMA, FP32 accumulation",
           "bf16_residual": "block-sparse BF16 MFMA with hi/lo residual correction, FP32 accumulation",
           "fp16": "block-sparse FP16 MFMA, power-of-two scales per ring and per row chunk, FP32 accumulation",
           "fp32": "FP32 CSR SpMM"}


if __name__ == "__main__":
    main()
