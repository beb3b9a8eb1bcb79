"""PF products on the matrix cores vs the FP64 oracle, for both storages of the PF: dense bf16 fragments
(pf_mode "dense") and 32 x 32 blocks of the non-zeros (pf_mode "sparse", one copy for both directions), and every
16-bit precision ("bf16", "bf16_residual", and "fp16" on the blocks).

Tolerances (relative to the largest entry of the oracle's result, and relative L2):
- bf16 (8-bit significand operands, FP32 accumulation): each product carries the rounding of both operands (about
  2.3e-3 standard deviation), an output summing n terms about 2.3e-3 / sqrt(n): L2 1e-3 where outputs sum many
  terms (K = 300, sigma 6 deg), 3e-3 on the tiny problems (a handful of terms per output), max 1e-2;
- fp16 (11-bit significands, the operands scaled by powers of two per ring and per row chunk, so no operand leaves
  fp16's normal range): unit roundoff 2^-11, a product about 4e-4 standard deviation: max 2e-3, L2 8e-4;
- bf16_residual (hi/lo operands, hi*hi + hi*lo + lo*hi): max 3e-5, L2 1e-5 (the FP32 path's bar is 2e-5 max).
"""
import numpy as np
import pytest

from pairs import build, cfg, rel_l2, rel_max

pytestmark = pytest.mark.gpu

TOL = {"bf16": (1e-2, 3e-3, 3e-3), "fp16": (2e-3, 8e-4, 8e-4), "bf16_residual": (3e-5, 1e-5, 1e-5)}  # max, L2, adj.
# (PF storage, precision): pf_mode "dense" = dense fragments, "sparse" = blocks of the non-zeros
CASES = [("dense", "bf16"), ("dense", "bf16_residual"), ("sparse", "bf16"), ("sparse", "bf16_residual"),
         ("sparse", "fp16"), ("generated", "bf16_residual"), ("generated", "fp16")]


def _inputs(op, seed=1):
    rng = np.random.default_rng(seed)
    x = (rng.random(op.coeff_shape) * op.support_mask()[None]).astype(np.float32)
    y = rng.random(op.data_shape).astype(np.float32)
    return x, y


@pytest.mark.parametrize("mode,precision", CASES)
@pytest.mark.parametrize("shape", ["small", "split_k", "batches", "tall"])
def test_forward_adjoint_vs_oracle(mode, precision, shape):
    print(f"case {mode} precision={precision} shape={shape}", flush=True)
    kw = dict(pf_mode=mode, pf_precision=precision)
    if shape == "small":            # one k chunk: direct write, no split
        op, orc, _ = build(cfg(), K=12, **kw)
    elif shape == "split_k":        # K = 300 in one batch: the k split and its reduction
        op, orc, _ = build(cfg(n_omega=12), K=300, sigma_deg=6.0, **kw)
    elif shape == "batches":        # 3 orientation batches, accumulated forward; C*P = 72 (not a tile multiple)
        c = cfg(nx=14, ny=12, my=17, n_omega=16, angle_range=(0, 360), cor=0.7, n_eta=24)
        op, orc, _ = build(c, K=70, sigma_deg=4.0, max_gb=3e-4, **kw)
        assert len(op.batches) > 1
    else:                           # My = 140: 9 y tiles in 2 groups
        c = cfg(nx=20, ny=20, my=140, n_omega=8, n_eta=20)
        op, orc, _ = build(c, K=40, **kw)
    assert {"dense": op.mfma_batches, "sparse": op.blk_batches, "generated": op.gen_batches}[mode] is not None
    tmax, tl2, tadj = TOL[precision]
    if precision == "bf16" and shape == "split_k":
        tl2 = 1e-3
    x, y = _inputs(op)
    d, d_ref = op.direct(x), orc.direct(x)
    a, a_ref = op.adjoint(y), orc.adjoint(y)
    print(f"  forward max {rel_max(d, d_ref):.2e} L2 {rel_l2(d, d_ref):.2e}; "
          f"adjoint max {rel_max(a, a_ref):.2e} L2 {rel_l2(a, a_ref):.2e}", flush=True)
    assert rel_max(d, d_ref) < tmax and rel_l2(d, d_ref) < tl2
    assert rel_max(a, a_ref) < tmax and rel_l2(a, a_ref) < tl2
    lhs = float(np.dot(x.ravel().astype(np.float64), a.ravel()))
    rhs = float(np.dot(d.ravel().astype(np.float64), y.ravel()))
    assert abs(lhs / rhs - 1) < tadj


@pytest.mark.parametrize("mode,precision", CASES)
def test_subsets(mode, precision):
    """Subset rows equal the full rows, the other rows are untouched, subset adjoints of a partition sum to the
    full adjoint (tests/test_subsets.py:49-86), on the matrix-core path."""
    import torch
    print(f"case {mode} precision={precision}", flush=True)
    c = cfg(nx=20, ny=20, my=21, n_omega=30, angle_range=(0, 360), cor=0.5)
    op, _, _ = build(c, K=90, pf_mode=mode, pf_precision=precision, max_gb=4e-4, oracle=False)
    assert len(op.batches) > 1
    subs = op.set_omega_subsets([np.arange(0, 7), np.arange(7, 20), np.arange(20, op.N_Omega)])
    x, y = _inputs(op, 3)
    xt = torch.from_numpy(x).cuda()
    yt = torch.from_numpy(y).cuda()
    full = torch.empty(op.data_shape, dtype=torch.float32, device="cuda")
    op.direct_cl(xt, full)
    adj_full = torch.empty(op.coeff_shape, dtype=torch.float32, device="cuda")
    op.adjoint_cl(yt, adj_full)
    acc = torch.zeros_like(adj_full)
    tmp = torch.empty_like(adj_full)
    for m, s in enumerate(subs):
        d = torch.full(op.data_shape, -7.0, dtype=torch.float32, device="cuda")
        op.direct_cl(xt, d, subset=m)
        rows = torch.from_numpy(s).cuda()
        assert torch.equal(d[rows], full[rows])
        mask = torch.ones(op.N_Omega, dtype=torch.bool, device="cuda")
        mask[rows] = False
        assert bool((d[mask] == -7.0).all())
        op.adjoint_cl(yt, tmp, subset=m)
        acc += tmp
    assert rel_max(acc.cpu().numpy(), adj_full.cpu().numpy()) < 1e-5


@pytest.mark.parametrize("mode,precision", CASES)
def test_spdhg_vs_oracle(mode, precision):
    """SPDHG epochs on the matrix-core path vs the FP64 oracle (same subset order from the same seed): with the
    residual correction the iterate is within 1e-4 (relative L2) after 6 epochs, as the FP32 path; fp16 within
    1e-3 (objectives 1e-4); plain bf16 within 3e-3, objectives within 1e-3."""
    from diffractom_amd import SPDHGPoisson
    from oracle.model import OracleSPDHG
    print(f"case spdhg {mode} {precision}", flush=True)
    op, orc, _ = build(cfg(nx=16, ny=16, my=17, n_omega=24, cor=0.0), K=40, sigma_deg=4.0, pf_mode=mode,
                       pf_precision=precision)
    rng = np.random.default_rng(7)
    xt = (rng.random(op.coeff_shape) * (rng.random(op.coeff_shape) < 0.2) * op.support_mask()[None]
          ).astype(np.float32)
    ax = orc.direct(xt)
    b = rng.poisson(ax / ax.mean() * 50).astype(np.float32)
    sp = SPDHGPoisson(op, subsets=4, seed=1)
    x = sp.run(np.zeros(op.coeff_shape, np.float32), b, 6, diagnostics_interval=3)
    ref = OracleSPDHG(orc, subsets=4, seed=1)
    x_ref = ref.run(np.zeros(op.coeff_shape), b, 6, diagnostics_interval=3)
    err = rel_l2(x, x_ref)
    f = [s["f"] for s in sp.iter_stats if s["f"] is not None]
    f_ref = [s["f"] for s in ref.iter_stats if s["f"] is not None]
    print(f"  x_6 rel L2 {err:.2e}; f {f} vs {f_ref}", flush=True)
    tx, tf = {"bf16_residual": (1e-4, 1e-5), "fp16": (1e-3, 1e-4), "bf16": (3e-3, 1e-3)}[precision]
    assert err < tx
    np.testing.assert_allclose(f, f_ref, rtol=tf)


def test_blocks_hold_the_csr_nonzeros():
    """The block storage holds exactly the non-zeros of the CSR forms (every block's count; the FP32 CSR and the
    blocks come from the same dense evaluation), and the fp16 values with their ring scales undone equal the FP32
    entries to fp16 rounding."""
    import torch
    print("case blocks vs csr", flush=True)
    c = cfg(nx=12, ny=12, my=13, n_omega=10, angle_range=(0, 180), cor=0.2, n_eta=40)
    op32, _, _ = build(c, K=75, pf_mode="sparse", oracle=False)
    for q in ("bf16_residual", "fp16"):
        op16, _, _ = build(c, K=75, pf_mode="sparse", pf_precision=q, oracle=False)
        assert op16.sparse_nnz == op32.sparse_nnz > 0
        blk = op16.blk_batches[0]
        assert int(blk["ptr"][-1].item()) == blk["nnz"] == op16.sparse_nnz
        # rebuild the dense PF from the blocks and compare with the FP32 CSR (forward rows (r, j) listing k)
        R, CP, kb = op16.N_Omega, op16.N_seg, op16.batches[0]["K_batch"]
        n_kb, n_jb = blk["n_kb"], blk["n_jb"]
        ptr = blk["ptr"].cpu().numpy().astype(np.int64)
        pos = blk["pos"].cpu().numpy().view(np.uint16).astype(np.int64)
        cnt = np.diff(ptr)
        bid = np.repeat(np.arange(len(cnt)), cnt)
        r, t = bid // (n_kb * n_jb), bid % (n_kb * n_jb)
        k = 32 * (t // n_jb) + (pos & 31)
        j = 32 * (t % n_jb) + (pos >> 5)
        if q == "fp16":
            v16 = blk["val"].cpu().numpy().view(np.float16).astype(np.float64)
            v = v16 / blk["ring_scale"].cpu().numpy().astype(np.float64)[j % op16.N_peaks]
        else:
            u = blk["val"].cpu().numpy().view(np.uint32)
            hi = ((u & 0xFFFF).astype(np.uint32) << 16).view(np.float32).astype(np.float64)
            lo = ((u >> 16).astype(np.uint32) << 16).view(np.float32).astype(np.float64)
            v = hi + lo
        dense = np.zeros((R, kb, CP))
        dense[r, k, j] = v
        sb = op32.sparse_batches[0]
        rp = sb["row_ptr_f"].cpu().numpy().astype(np.int64)
        ck = sb["col_k"].cpu().numpy().view(np.uint16).astype(np.int64)
        vf = sb["val_f"].cpu().numpy().astype(np.float64)
        rows = np.repeat(np.arange(R * CP), np.diff(rp))
        ref = np.zeros((R, kb, CP))
        ref[rows // CP, ck, rows % CP] = vf
        assert np.array_equal(dense != 0, ref != 0)
        tol = 2.0 ** -11 if q == "fp16" else 2.0 ** -16
        assert np.all(np.abs(dense - ref) <= tol * np.abs(ref)), q


@pytest.mark.parametrize("precision", ["bf16_residual", "fp16"])
def test_generated_blocks_equal_stored_blocks(precision):
    """The PF generated per call (dfm_pfgen: candidate eta windows, then the stored mode's entry function) holds the
    stored blocks' entries: the same non-zero count, subset products bit for bit equal to the stored mode's (the same
    fragments), the full products (generated in angle chunks, so their k splits differ) within 1e-6; three
    orientation batches, Gauss-Legendre eta nodes, an eta range of half a turn."""
    import torch
    print(f"case generated vs stored {precision}", flush=True)
    from diffractom_amd import QuadratureRule
    c = cfg(nx=14, ny=12, my=17, n_omega=20, angle_range=(0, 360), cor=0.7, n_eta=48)
    c["eta_angle_range"] = [0, 180]
    kw = dict(K=70, sigma_deg=3.0, max_gb=3e-4, pf_precision=precision, oracle=False,
              quadrature=QuadratureRule(1, 1, 2, 1))
    ops = {m: build(c, pf_mode=m, **kw)[0] for m in ("sparse", "generated")}
    assert len(ops["sparse"].batches) > 1 and ops["generated"].gen_batches is not None
    assert ops["sparse"].sparse_nnz == ops["generated"].sparse_nnz > 0
    x, y = _inputs(ops["sparse"], 4)
    xt, yt = torch.from_numpy(x).cuda(), torch.from_numpy(y).cuda()
    res = {}
    for m, op in ops.items():
        op.set_omega_subsets(4)
        out = []
        for sub in (1, 3, None):
            d = torch.zeros(op.data_shape, device="cuda")
            a = torch.zeros(op.coeff_shape, device="cuda")
            op.direct_cl(xt, d, subset=sub)
            op.adjoint_cl(yt, a, subset=sub)
            out.append((d.cpu().numpy(), a.cpu().numpy()))
        res[m] = out
    for i, sub in enumerate((1, 3, None)):
        for u, v in zip(res["generated"][i], res["sparse"][i]):
            if sub is None:
                assert rel_max(u, v) < 1e-6
            else:
                assert np.array_equal(u, v)


def test_generated_slots_follow_the_angle_group():
    """The generated blocks stay in their batch slots for the next product on the same subset and are generated
    again when another angle group (another subset, a full product's angle chunks, a new subset partition) took the
    slots: every product in such a sequence equals the stored mode's bit for bit."""
    import torch
    print("case generated slot reuse", flush=True)
    c = cfg(nx=14, ny=12, my=17, n_omega=20, angle_range=(0, 360), cor=0.7)
    kw = dict(K=70, sigma_deg=3.0, max_gb=3e-4, pf_precision="fp16", oracle=False)
    ops = {m: build(c, pf_mode=m, **kw)[0] for m in ("sparse", "generated")}
    x, y = _inputs(ops["sparse"], 5)
    xt, yt = torch.from_numpy(x).cuda(), torch.from_numpy(y).cuda()
    seq = [("set", 4), ("f", 1), ("a", 1), ("f", 3), ("a", 1), ("f", None), ("a", 3), ("a", 3),
           ("set", [np.arange(0, 20, 2), np.arange(1, 20, 2)]), ("a", 1), ("f", 0), ("a", 0)]
    res = {}
    for m, op in ops.items():
        out = []
        for what, sub in seq:
            if what == "set":
                op.set_omega_subsets(sub)
                continue
            if what == "f":
                d = torch.zeros(op.data_shape, device="cuda")
                op.direct_cl(xt, d, subset=sub)
                out.append(d.cpu().numpy())
            else:
                a = torch.zeros(op.coeff_shape, device="cuda")
                op.adjoint_cl(yt, a, subset=sub)
                out.append(a.cpu().numpy())
        res[m] = out
    for i, (u, v) in enumerate(zip(res["generated"], res["sparse"])):
        if i == 4:                   # the full forward: generated in angle chunks, so its k split differs
            assert rel_max(u, v) < 1e-6
        else:
            assert np.array_equal(u, v), i


@pytest.mark.parametrize("mode,precision", [("sparse", "bf16_residual"), ("sparse", "fp16"), ("generated", "fp16")])
def test_block_workgroup_sizes_agree(mode, precision):
    """The block-sparse products in 8-wave workgroups (eight column blocks share each staged y-side chunk) give the
    4-wave results bit for bit: every wave still owns one column block and walks the same chunks in the same order."""
    import torch
    print(f"case block workgroups {mode} {precision}", flush=True)
    c = cfg(nx=20, ny=20, my=41, n_omega=12, angle_range=(0, 180), cor=0.3)
    op, _, _ = build(c, K=300, pf_mode=mode, pf_precision=precision, oracle=False)
    op.set_omega_subsets(3)
    x, y = _inputs(op, 6)
    xt, yt = torch.from_numpy(x).cuda(), torch.from_numpy(y).cuda()
    res = []
    for wg in (4, 8):
        op.blk_variant.update(fwd_wg=wg, adj_wg=wg)
        out = []
        for sub in (None, 1):
            d = torch.zeros(op.data_shape, device="cuda")
            a = torch.zeros(op.coeff_shape, device="cuda")
            op.direct_cl(xt, d, subset=sub)
            op.adjoint_cl(yt, a, subset=sub)
            out += [d.cpu().numpy(), a.cpu().numpy()]
        res.append(out)
    for u, v in zip(*res):
        assert np.array_equal(u, v)


@pytest.mark.parametrize("my", [103, 90])
@pytest.mark.parametrize("mode,precision", [("sparse", "bf16_residual"), ("sparse", "fp16"), ("generated", "fp16")])
def test_block_fixed_tile_count_agrees(mode, precision, my):
    """The 4-wave block products with the y-tile count compiled in (7 tiles at 103 detector rows, 6 at 90) give the
    run-time count's results bit for bit (the same tiles, chunks and accumulation order)."""
    import torch
    print(f"case block fixed tile count {mode} {precision} my={my}", flush=True)
    c = cfg(nx=12, ny=12, my=my, n_omega=12, angle_range=(0, 180), cor=0.3)
    op, _, _ = build(c, K=300, pf_mode=mode, pf_precision=precision, oracle=False)
    op.set_omega_subsets(3)
    x, y = _inputs(op, 7)
    xt, yt = torch.from_numpy(x).cuda(), torch.from_numpy(y).cuda()
    res = []
    for fixed in (False, True):
        op.blk_variant.update(fwd_wg=4, adj_wg=4, fwd_fixed=fixed, adj_fixed=fixed)
        out = []
        for sub in (None, 1):
            d = torch.zeros(op.data_shape, device="cuda")
            a = torch.zeros(op.coeff_shape, device="cuda")
            op.direct_cl(xt, d, subset=sub)
            op.adjoint_cl(yt, a, subset=sub)
            out += [d.cpu().numpy(), a.cpu().numpy()]
        res.append(out)
    for u, v in zip(*res):
        assert np.array_equal(u, v)
