"""HIP operator vs the FP64 oracle: forward, adjoint, transpose, subsets, storage forms, power iteration.
Tolerances: FP32 forward/adjoint within 2e-5 of the largest entry (the reference's own GPU-vs-FP64 bar,
tests/test_differentiable.py:36-41); adjointness 1e-5 (tests/test_adjoint.py:43-47)."""
import numpy as np
import pytest

from pairs import build, cfg, rel_max

pytestmark = pytest.mark.gpu


def _inputs(op, seed=1):
    rng = np.random.default_rng(seed)
    x = (rng.random(op.coeff_shape) * op.support_mask()[None]).astype(np.float32)
    y = rng.random(op.data_shape).astype(np.float32)
    return x, y


@pytest.mark.parametrize("angle_range", [(0, 180), (0, 360), (0, 90), (180, 0)])
@pytest.mark.parametrize("kernel", ["trapezoid", "hat"])
def test_forward_adjoint_match_oracle(angle_range, kernel):
    print(f"case angle_range={angle_range} kernel={kernel}", flush=True)
    op, orc, _ = build(cfg(angle_range=angle_range), radon_kernel=kernel, pf_mode="sparse")
    x, y = _inputs(op)
    d = op.direct(x)
    d_ref = orc.direct(x)
    assert rel_max(d, d_ref) < 2e-5
    a = op.adjoint(y)
    a_ref = orc.adjoint(y)
    assert rel_max(a, a_ref) < 2e-5
    lhs = float(np.dot(x.ravel().astype(np.float64), a.ravel()))
    rhs = float(np.dot(d.ravel().astype(np.float64), y.ravel()))
    assert rhs > 0 and abs(lhs / rhs - 1) < 1e-5


@pytest.mark.parametrize("pf_mode", ["sparse", "dense", "generated"])
def test_storage_forms_and_rectangular_grid(pf_mode):
    """Nx != Ny, a centre-of-rotation offset, small kernels (Gauss-Legendre quadrature), 4 orientation batches."""
    print(f"case pf_mode={pf_mode}", flush=True)
    c = cfg(nx=14, ny=12, my=17, n_omega=16, angle_range=(0, 360), cor=0.7, n_eta=24)
    op, orc, _ = build(c, K=12, sigma_deg=1.0, pf_mode=pf_mode, max_gb=3e-5)
    assert len(op.batches) > 1
    assert not op.quadrature.is_midpoint
    x, y = _inputs(op, 2)
    assert rel_max(op.direct(x), orc.direct(x)) < 2e-5
    assert rel_max(op.adjoint(y), orc.adjoint(y)) < 2e-5


@pytest.mark.parametrize("pf_mode", ["sparse", "generated", "dense"])
@pytest.mark.parametrize("subsets", [4, "blocks"])
def test_subset_forward_and_adjoint(pf_mode, subsets):
    """tests/test_subsets.py:49-86: subset rows equal the full rows, other rows untouched, each subset adjoint is
    the transpose of its forward, the subset adjoints of a partition sum to the full adjoint."""
    import torch
    print(f"case pf_mode={pf_mode} subsets={subsets}", flush=True)
    c = cfg(nx=20, ny=20, my=21, n_omega=30, angle_range=(0, 360), cor=0.5)
    op, orc, _ = build(c, K=30, pf_mode=pf_mode, max_gb=1e-4, oracle=True)
    assert len(op.batches) > 1
    R = op.N_Omega
    subs = op.set_omega_subsets(4 if subsets == 4 else [np.arange(0, 7), np.arange(7, 20), np.arange(20, R)])
    x, y = _inputs(op, 6)
    xd, yd = torch.from_numpy(x).cuda(), torch.from_numpy(y).cuda()
    full = op.direct(x)
    adj = op.adjoint(y)
    adj_sum = np.zeros_like(adj, dtype=np.float64)
    for m, s in enumerate(subs):
        d = torch.full(op.data_shape, -7.0, dtype=torch.float32, device="cuda")
        op.direct_cl(xd, d, subset=m)
        d = d.cpu().numpy()
        np.testing.assert_allclose(d[s], full[s], rtol=1e-5, atol=1e-6 * np.abs(full).max())
        rest = np.setdiff1d(np.arange(R), s)
        assert np.all(d[rest] == -7.0)
        a_m = torch.empty(op.coeff_shape, dtype=torch.float32, device="cuda")
        op.adjoint_cl(yd, a_m, subset=m)
        a_m = a_m.cpu().numpy()
        adj_sum += a_m
        lhs = float(np.dot(d[s].ravel().astype(np.float64), y[s].ravel()))
        rhs = float(np.dot(x.ravel().astype(np.float64), a_m.ravel()))
        assert abs(lhs / rhs - 1) < 1e-5
    np.testing.assert_allclose(adj_sum, adj, rtol=1e-4, atol=1e-5 * np.abs(adj).max())
    orc.set_omega_subsets(subs)
    assert rel_max(op.adjoint(np.where(np.isin(np.arange(R), subs[0])[:, None, None], y, 0)),
                   orc.adjoint(y, subset=0)) < 2e-5


def test_dense_fp32_subsets():
    """The reference raises NotImplementedError for subsets of a dense-PF operator
    (single_phase_forward_operator.py:1564-1566); here the dense FP32 path supports them: subset rows equal the
    full rows."""
    import torch
    op, _, _ = build(pf_mode="dense", oracle=False)
    subs = op.set_omega_subsets(2)
    x, _ = _inputs(op)
    xt = torch.from_numpy(x).cuda()
    full = torch.empty(op.data_shape, dtype=torch.float32, device="cuda")
    op.direct_cl(xt, full)
    d = torch.zeros_like(full)
    op.direct_cl(xt, d, subset=1)
    rows = torch.from_numpy(subs[1]).cuda()
    assert torch.allclose(d[rows], full[rows], rtol=1e-6, atol=1e-6 * float(full.abs().max()))


def test_power_iteration_matches_oracle():
    from diffractom_amd import estimate_L_power
    from oracle.model import estimate_L_power as oracle_power
    op, orc, _ = build(pf_mode="sparse")
    L = estimate_L_power(op, niter=15, seed=3, verbose=0)
    L_ref = oracle_power(orc, niter=15, seed=3)
    assert abs(L / L_ref - 1) < 1e-5


def test_errors_raise_like_the_reference():
    import torch
    op, _, _ = build(pf_mode="sparse", oracle=False)
    with pytest.raises(TypeError):
        op.direct_cl(torch.zeros(op.coeff_shape, dtype=torch.float64, device="cuda"),
                     torch.zeros(op.data_shape, device="cuda"))
    with pytest.raises(ValueError):
        op.direct_cl(torch.zeros((1, 2, 3), device="cuda"), torch.zeros(op.data_shape, device="cuda"))
    with pytest.raises(ValueError):
        op.set_omega_subsets(0)
    with pytest.raises(ValueError):
        op.direct_cl(torch.zeros(op.coeff_shape, device="cuda"), torch.zeros(op.data_shape, device="cuda"), subset=0)


@pytest.mark.parametrize("kernel,channels", [("table", 4), ("table_unr4", 4), ("table_band", 4), ("table_ch128", 4),
                                             ("table_band128", 4), ("table_band256", 4),
                                             ("table_group", 4), ("table_tile8", 4), ("table_tile8g2", 4),
                                             ("sweep", 4), ("sweep2", 4), ("sweep32", 4), ("sweep32_1", 4),
                                             ("psweep1", 4), ("psweep2", 4), ("psweep4", 4), ("psweep1np", 4),
                                             ("psweep2np", 4), ("psweep1dup", 4), ("psweep2dup", 4), ("arith", 4)])
@pytest.mark.parametrize("shape", ["square", "chunked"])
def test_projector_kernels_match_oracle(kernel, channels, shape):
    """Every projector kernel choice (geometry tables, staged or not, capped; pixel bands; line sweeps with one or two
    rays per thread; the backprojection's tile, grouped and 8x8 forms; footprint arithmetic) against the
    oracle; "chunked": 120 angles x 103 bins (several table chunks of angles)."""
    print(f"case kernel={kernel} channels={channels} shape={shape}", flush=True)
    if shape == "square":
        c = cfg(nx=14, ny=12, my=17, n_omega=16, angle_range=(0, 360), cor=0.7, n_eta=24)
    else:
        c = cfg(nx=16, ny=16, my=103, n_omega=120, angle_range=(0, 180), cor=0.3, n_eta=12)
    op, orc, _ = build(c, K=8, pf_mode="sparse")
    op.radon_variant["fwd_kernel"] = op.radon_variant["bwd_kernel"] = kernel.split("_")[0]
    if kernel == "table_unr4":                   # 4 table entries ahead (default 8)
        op.radon_variant["fwd"] = 4 + 256
    elif kernel.startswith("table_band"):        # forward tiled by 3 pixel bands, band partial sums
        op.radon_bands = 3
        op._tables_full = op._radon_tables(np.arange(op.N_Omega))
        op.radon_variant["fwd_kernel"] = "band"
        op.radon_variant["fwd_band"] = 8 + 16 * {"table_band": 0, "table_band128": 1, "table_band256": 2}[kernel]
    elif kernel.startswith("sweep"):             # forward by line sweep (4 or 2 lines per LDS block)
        op.radon_variant["fwd_kernel"] = "sweep"
        op.radon_variant["fwd_sweep"] = {"sweep": 4, "sweep2": 2, "sweep32": 2 + 16, "sweep32_1": 1 + 16}[kernel]
        op.radon_variant["bwd_kernel"] = "table"
    elif kernel.startswith("psweep"):            # forward by line sweep with two rays per thread
        op.radon_variant["fwd_kernel"] = "psweep"
        # (np: each staged block loaded at its start, two workgroups per CU; dup: the duplicated-pixel LDS layout)
        op.radon_variant["fwd_psweep"] = int(kernel[6]) + {"np": 8, "up": 16}.get(kernel[-2:], 0)
        op.radon_variant["bwd_kernel"] = "table"
        assert op._tables_full["sweep_grp_p"][0, 0] >= 0
    elif kernel == "table_group":                # backprojection: a workgroup per tile and 4 channel blocks
        op.radon_variant["bwd"] = 8 + 64 * 4
    elif kernel == "table_tile8":                # backprojection over 8x8 tiles, 4 angles per pass
        op.radon_variant["bwd"] = 4 + 64 * 1 + 1024
    elif kernel == "table_tile8g2":              # ... 8 angles per pass, 2 channel blocks
        op.radon_variant["bwd"] = 8 + 64 * 2 + 1024
    elif kernel == "table_ch128":                # forward channel blocks of 128 channels
        op.radon_variant["fwd"] = 8 + 256 + 512 * 3
    x, y = _inputs(op, 5)
    assert rel_max(op.direct(x), orc.direct(x)) < 2e-5
    assert rel_max(op.adjoint(y), orc.adjoint(y)) < 2e-5


@pytest.mark.parametrize("variant", [8 + 64 * 4, 4 + 64 * 2, 8 + 64 * 8, 4 + 64 + 1024, 8 + 64 * 2 + 1024,
                                     8 + 64 * 4 + 1024])
def test_backprojection_variants_agree(variant):
    """Every grouped backprojection form (4x4 tiles with G channel blocks of 64 channels; 8x8 tiles) gives the
    one-tile-per-workgroup kernel's values (pinned against the oracle by the
    adjoint tests), plain (mode 0) and with the SPDHG primal update (mode 1: z += A^T y, x = max(0, x - T (z + A^T y
    / p + lam)) where T > 0, else 0): K = 300 orientations (several blocks, the last one partial), one subset of
    angles; relative max difference within 1e-6 (the same angle order per pixel)."""
    import torch
    print(f"case backprojection variant={variant}", flush=True)
    c = cfg(nx=16, ny=13, my=23, n_omega=20, angle_range=(0, 180), cor=0.3, n_eta=12)
    op, _, _ = build(c, K=300, pf_mode="sparse", oracle=False)
    op.set_omega_subsets(2)
    n_ang, _, _, geo_b = op._angle_set(1)
    tab = op._tables(1)
    if (variant // 1024) % 4 == 1 and tab["tbase8"].numel() == 1:
        pytest.skip("8x8 tiles reach more than 16 bins for this geometry")
    b = op.batches[0]
    ks, kn = op.K_stride, b["kstride"]
    npix = op.Nx * op.Ny
    g = torch.Generator(device="cuda")
    g.manual_seed(3)
    sino = torch.rand(n_ang * op.My * kn, generator=g, device="cuda")
    x0 = torch.rand(npix * ks, generator=g, device="cuda")
    z0 = torch.rand(npix * ks, generator=g, device="cuda") - 0.5
    T0 = torch.rand(npix * ks, generator=g, device="cuda") * (torch.rand(npix * ks, generator=g, device="cuda") > 0.1)
    res = {}
    for v in (4, variant):
        op.radon_variant["bwd_kernel"] = "table"
        op.radon_variant["bwd"] = v
        img = torch.zeros(npix * ks, device="cuda")
        op._radon_backward(sino, geo_b, tab, n_ang, img, ks, b["koff"], kn)
        x, z = x0.clone(), z0.clone()
        op._radon_backward(sino, geo_b, tab, n_ang, None, ks, b["koff"], kn, x=x, z=z, T=T0, inv_p=0.7, lam=0.05)
        torch.cuda.synchronize()
        res[v] = (img.cpu().numpy(), x.cpu().numpy(), z.cpu().numpy())
    for a, r in zip(res[variant], res[4]):
        assert rel_max(a, r) < 1e-6


@pytest.mark.parametrize("precision", ["fp32", "bf16_residual"])
def test_autotune_keeps_results(precision):
    """The autotuner times every candidate of each product and keeps the fastest; the operator then still matches the
    oracle (every candidate computes the same values)."""
    print(f"case autotune {precision}", flush=True)
    c = cfg(nx=20, ny=20, my=21, n_omega=30, angle_range=(0, 180), cor=0.5)
    kw = {"pf_mode": "sparse"} if precision == "fp32" else {"pf_mode": "dense", "pf_precision": precision}
    op, orc, _ = build(c, K=24, **kw)
    op.set_omega_subsets(5)
    tuning = op.autotune(subset=1, reps=2)
    assert "radon_forward" in tuning and "radon_backward_spdhg" in tuning
    if precision != "fp32":
        assert "pf_forward" in tuning and "pf_adjoint" in tuning
    for name, (choice, times) in tuning.items():
        assert choice in times and all(t > 0 for t in times.values())
    x, y = _inputs(op, 7)
    tol = 2e-5 if precision == "fp32" else 3e-5
    assert rel_max(op.direct(x), orc.direct(x)) < tol
    assert rel_max(op.adjoint(y), orc.adjoint(y)) < tol


@pytest.mark.parametrize("precision", ["fp32", "bf16_residual"])
def test_hexagonal_titanium_matches_oracle(precision):
    """The C2 material: Ti (P6_3/mmc, a 2.95 A, c 4.68 A, |G| = 12), the 8 rings {10-10} ... {11-22} (planes as
    three-index hkl), 360 degrees of omega, sigma 2.5 deg; forward and adjoint against the oracle (hexagonal pole
    axes, the a* || x frame of the non-orthogonal cell)."""
    from diffractom_amd import MatrixGrid, SinglePhaseForwardOperator
    from diffractom_amd.workloads import TI_RINGS, random_fz, ti_material
    from oracle.model import OracleOperator
    print(f"case Ti {precision}", flush=True)
    c = cfg(nx=16, ny=16, my=17, n_omega=24, angle_range=(0, 360), cor=0.4, n_eta=36, wavelength=0.2)
    U = random_fz(10, "hexagonal", 4)
    sigma = np.deg2rad(2.5)
    kw = {"pf_mode": "sparse"} if precision == "fp32" else {"pf_mode": "dense", "pf_precision": precision}
    op = SinglePhaseForwardOperator(cfg=c, material=ti_material(0.2), grid=MatrixGrid(U, sigma), max_gb=0.2,
                                    normalized=True, **kw)
    orc = OracleOperator(c, (2.95, 2.95, 4.68, 90, 90, 120), "hexagonal", TI_RINGS, 0.2, U, sigma,
                         orthonormalize=False)
    assert op.N_peaks == orc.P == 8 and op.data_shape == orc.data_shape
    x, y = _inputs(op, 9)
    tol = 2e-5 if precision == "fp32" else 3e-5
    assert rel_max(op.direct(x), orc.direct(x)) < tol
    assert rel_max(op.adjoint(y), orc.adjoint(y)) < tol


@pytest.mark.parametrize("ksplit,kc", [(1, 768), (3, 32), (8, 256)])
def test_csr_forward_split_matches_oracle(ksplit, kc):
    """The row-per-thread CSR forward with the orientations split over workgroup ranges (bisection to each range's
    first entry, partial sums reduced in a fixed order) and LDS passes of kc orientations; several batches."""
    print(f"case csr ksplit={ksplit} kc={kc}", flush=True)
    c = cfg(nx=14, ny=12, my=17, n_omega=16, angle_range=(0, 360), cor=0.7, n_eta=24)
    op, orc, _ = build(c, K=150, sigma_deg=5.0, pf_mode="sparse", max_gb=2.2e-4)
    assert len(op.batches) > 1
    op.csr_variant.update(fwd="rows", ksplit=ksplit, kc=kc)
    x, _ = _inputs(op, 4)
    assert rel_max(op.direct(x), orc.direct(x)) < 2e-5
    op.set_omega_subsets(3)
    import torch
    xt = torch.from_numpy(x).cuda()
    d = torch.zeros(op.data_shape, dtype=torch.float32, device="cuda")
    op.direct_cl(xt, d, subset=1)
    rows = op.omega_subsets[1]
    assert rel_max(d.cpu().numpy()[rows], orc.direct(x)[rows]) < 2e-5


@pytest.mark.parametrize("my,tsplit,tkc,direct", [(17, 1, 0, 3), (17, 4, 8, 3), (130, 3, 0, 3), (200, 2, 16, 3),
                                                  (300, 5, 0, 3), (17, 4, 8, 0), (130, 3, 0, 0), (300, 5, 0, 0)])
def test_csr_transposed_kernels_match_oracle(my, tsplit, tkc, direct):
    """The wavefront-segmented CSR products on transposed operands, both directions: 1, 3 and 5 groups of 64 detector
    positions per wave (16 or 8 rows per wave), orientation splits with partial sums, LDS chunks of tkc columns (0:
    as many as fit; 8 and 16 force many chunk and window changes); the sinogram transposed in the kernels (direct 3:
    staged through LDS by the forward, written through LDS by the adjoint) or by separate transposes (0); several
    batches, full and subset rows."""
    print(f"case csr transposed my={my} tsplit={tsplit} tkc={tkc} direct={direct}", flush=True)
    c = cfg(nx=12, ny=12, my=my, n_omega=12, angle_range=(0, 360), cor=0.7, n_eta=24)
    op, orc, _ = build(c, K=150, sigma_deg=5.0, pf_mode="sparse", max_gb=2.2e-4)
    assert len(op.batches) > 1
    op.csr_variant.update(fwd="t", adj="t", tsplit=tsplit, tkc=tkc, tdirect=direct)
    x, y = _inputs(op, 4)
    assert rel_max(op.direct(x), orc.direct(x)) < 2e-5
    assert rel_max(op.adjoint(y), orc.adjoint(y)) < 2e-5
    op.set_omega_subsets(3)
    import torch
    xt = torch.from_numpy(x).cuda()
    d = torch.zeros(op.data_shape, dtype=torch.float32, device="cuda")
    op.direct_cl(xt, d, subset=2)
    rows = op.omega_subsets[2]
    assert rel_max(d.cpu().numpy()[rows], orc.direct(x)[rows]) < 2e-5
    ysub = np.zeros_like(y)
    ysub[rows] = y[rows]
    out = torch.empty(op.coeff_shape, dtype=torch.float32, device="cuda")
    op.adjoint_cl(torch.from_numpy(y).cuda(), out, subset=2)
    assert rel_max(out.cpu().numpy(), orc.adjoint(ysub)) < 2e-5


def test_dense_fp32_reduction_split():
    """The FP32 dense forward with its reduction over the orientations split in parts summed in order (taken at
    large K; forced here with parts of 32 terms): the oracle's forward within 2e-5, subsets included."""
    print("case dense fp32 split", flush=True)
    op, orc, _ = build(cfg(n_omega=12), K=300, sigma_deg=6.0, pf_mode="dense")
    op.dense_split_min = 32
    x, y = _inputs(op, 4)
    assert rel_max(op.direct(x), orc.direct(x)) < 2e-5
    op.set_omega_subsets(3)
    orc.set_omega_subsets(3)
    import torch
    d = torch.zeros(op.data_shape, dtype=torch.float32, device="cuda")
    op.direct_cl(torch.from_numpy(x).cuda(), d, subset=1)
    assert rel_max(d.cpu().numpy(), orc.direct(x, subset=1)) < 2e-5


@pytest.mark.parametrize("nx", [99, 150])
def test_pairs_sweep_forms_agree_on_long_lines(nx):
    """The ray-pair sweep's forms (1, 2 or 4 staged lines; each block loaded at its start; every pixel held twice in
    LDS) compute every ray's taps in the same order, so their sinograms agree bit for bit; at 99 and 150 pixels a
    line (the 128-pixel limit of the four-line and the two-line loaded-at-start and duplicated forms in between),
    64 orientations,
    a subset of the angles; and one of them against the table forward (the projector the tests above pin)."""
    import torch
    print(f"case pairs sweep forms nx={nx}", flush=True)
    c = cfg(nx=nx, ny=nx, my=int(nx * 1.05) + 3, n_omega=20, angle_range=(0, 180), cor=0.4, n_eta=12)
    op, _, _ = build(c, K=64, pf_mode="sparse", oracle=False)
    op.set_omega_subsets(2)
    tab = op._tables(1)
    assert tab["pairs_ok"]
    n_ang, _, geo_f, _ = op._angle_set(1)
    b = op.batches[0]
    ks, kn = op.K_stride, b["kstride"]
    g = torch.Generator(device="cuda")
    g.manual_seed(11)
    img = torch.rand(op.Nx * op.Ny * ks, generator=g, device="cuda")
    forms = [1, 2, 9, 17] + ([4, 10, 18] if nx <= 128 else [])
    res = {}
    for v in forms + ["table"]:
        sino = torch.zeros(n_ang * op.My * kn, device="cuda")
        if v == "table":
            op.radon_variant.update(fwd_kernel="table", fwd=8 + 256)
        else:
            op.radon_variant.update(fwd_kernel="psweep", fwd_psweep=v)
        op._radon_forward(img, ks, b["koff"], kn, sino, geo_f, n_ang, tab)
        torch.cuda.synchronize()
        res[v] = sino.cpu().numpy()
    for v in forms[1:]:
        assert np.array_equal(res[v], res[forms[0]]), v
    assert rel_max(res[2], res["table"]) < 2e-5
