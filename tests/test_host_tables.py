"""The projector's geometry tables (host, float64) reproduce the oracle's projection matrices: the forward CSR is
scale/A K(t - s) of every pixel and bin, the pixel table the same weights times the backprojection weight pi/R, and
every pixel's two bins lie in the 8 bins of its tile (CPU)."""
import numpy as np
import pytest

from diffractom_amd.operator import radon_tables
from oracle import model as O
from pairs import AL3, cfg, orientations


def _geo(op, kernel):
    X, Y, T0 = op.X, op.Y, op.T0
    A = np.maximum(np.abs(X.astype(np.float32)), np.abs(Y.astype(np.float32))).astype(np.float64)
    if kernel == "hat":
        hw, top, inv_a = np.ones_like(A), np.ones_like(A), np.ones_like(A)
    else:
        hw, top, inv_a = 0.5 * (1 + A), np.minimum(1, A), 1 / A
    return dict(X=X, Y=Y, T0=T0, hw=hw, top=top, inv_a=inv_a, w=np.full(op.R, np.pi / op.R))


@pytest.mark.parametrize("kernel", ["trapezoid", "hat"])
@pytest.mark.parametrize("shape", [(16, 16, 17, (0, 180)), (14, 12, 17, (0, 360)), (99, 99, 103, (0, 180))])
def test_tables_equal_projection_matrices(kernel, shape):
    nx, ny, my, ar = shape
    op = O.OracleOperator(cfg(nx=nx, ny=ny, my=my, n_omega=30, angle_range=ar, cor=0.7), (4.05,) * 3 + (90,) * 3,
                          "cubic", AL3, 0.35, orientations(2, 1), np.deg2rad(4.0), radon_kernel=kernel,
                          orthonormalize=False)
    idx = np.array([0, 3, 7, 11, 14, 22, 29])
    tab = radon_tables(_geo(op, kernel), idx, nx, ny, my, op.scale)
    npix = nx * ny
    for i, r in enumerate(idx):
        P = op.radon_matrix(r).toarray()                                   # (My, Npix) with scale
        F = np.zeros_like(P)
        lo, hi = tab["row_ptr"][i * my], tab["row_ptr"][(i + 1) * my]
        rows = np.repeat(np.arange(my), np.diff(tab["row_ptr"][i * my:(i + 1) * my + 1]))
        F[rows, tab["pix"][lo:hi]] = tab["w"][lo:hi]
        A = max(abs(np.float32(op.X[r])), abs(np.float32(op.Y[r])))
        # the forward table carries 1/A and the scale of the trapezoid; the oracle's K already holds 1/A
        np.testing.assert_allclose(F, P, rtol=2e-6, atol=1e-7 * P.max())
        B = np.zeros_like(P)
        pt = tab["ptab"][i]
        w0, w1 = pt[:, 2].view(np.float32), pt[:, 3].view(np.float32)
        np.add.at(B, (pt[:, 0], np.arange(npix)), w0)
        np.add.at(B, (pt[:, 1], np.arange(npix)), w1)
        np.testing.assert_allclose(B, P / op.scale * np.pi / op.R, rtol=2e-6, atol=1e-7 * B.max())
        tnx = -(-nx // 4)
        yy, xx = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
        tile = (yy.ravel() // 4) * tnx + xx.ravel() // 4
        base = tab["tbase"][i][tile]
        assert np.all((pt[:, 0] - base >= 0) & (pt[:, 0] - base < 8) & (pt[:, 1] - base >= 0) & (pt[:, 1] - base < 8))
        tile8 = (yy.ravel() // 8) * -(-nx // 8) + xx.ravel() // 8       # 8 x 8 tiles: 16 bins from their base
        base8 = tab["tbase8"][i][tile8]
        assert np.all((pt[:, :2] - base8[:, None] >= 0) & (pt[:, :2] - base8[:, None] < 16))


@pytest.mark.parametrize("n_bands", [1, 3, 5])
def test_band_pointers_split_each_ray_by_image_rows(n_bands):
    """band_ptr cuts every ray's entries into n_bands contiguous ranges, range b holding exactly the entries whose
    pixel lies in image rows [b h, (b + 1) h), h = ceil(ny / n_bands): the banded forward sums the same products."""
    nx, ny, my = 14, 12, 17
    op = O.OracleOperator(cfg(nx=nx, ny=ny, my=my, n_omega=16, angle_range=(0, 360), cor=0.7),
                          (4.05,) * 3 + (90,) * 3, "cubic", AL3, 0.35, orientations(2, 1), np.deg2rad(4.0),
                          orthonormalize=False)
    idx = np.array([0, 3, 7, 11])
    tab = radon_tables(_geo(op, "trapezoid"), idx, nx, ny, my, op.scale, n_bands)
    h = -(-ny // n_bands)
    bp, rp = tab["band_ptr"], tab["row_ptr"]
    assert len(bp) == len(idx) * my * n_bands + 1
    for r in range(len(idx) * my):
        assert bp[r * n_bands] == rp[r] and bp[(r + 1) * n_bands] == rp[r + 1]
        for b in range(n_bands):
            rows = tab["pix"][bp[r * n_bands + b]:bp[r * n_bands + b + 1]] // nx
            assert np.all((rows >= b * h) & (rows < (b + 1) * h))


@pytest.mark.parametrize("shape", [(14, 12, 17, (0, 360)), (16, 16, 17, (0, 180))])
def test_sweep_geometry_reproduces_the_table(shape):
    """The line sweep's rule (pixels of line L with |coef p + cross L + T0 - s| < hw, weight scale * min(hw - |d|,
    top); rows for |X| >= |Y|, columns otherwise) lists exactly the table's pixels and weights of every ray, and every
    angle sits in one group of its class with at most 1024 rays per group."""
    nx, ny, my, ar = shape
    op = O.OracleOperator(cfg(nx=nx, ny=ny, my=my, n_omega=16, angle_range=ar, cor=0.7), (4.05,) * 3 + (90,) * 3,
                          "cubic", AL3, 0.35, orientations(2, 1), np.deg2rad(4.0), orthonormalize=False)
    idx = np.arange(16)
    g = _geo(op, "trapezoid")
    tab = radon_tables(g, idx, nx, ny, my, op.scale)
    rp, geo = tab["row_ptr"], tab["sweep_geo"]
    cls = (np.abs(g["X"][idx]) < np.abs(g["Y"][idx])).astype(int)
    for r in range(len(idx) * my):
        a, s = divmod(r, my)
        X, Y, T0, hw, top, scale = geo[a]
        coef, cross, ln, nl = (Y, X, ny, nx) if cls[a] else (X, Y, nx, ny)
        pix, wts = [], []
        for L in range(nl):
            c = cross * L + T0 - s
            e0, e1 = (-hw - c) / coef, (hw - c) / coef
            for p in range(max(int(np.ceil(min(e0, e1))), 0), min(int(np.floor(max(e0, e1))), ln - 1) + 1):
                k = hw - abs(coef * p + c)
                if k > 0:
                    pix.append(p * nx + L if cls[a] else L * nx + p)
                    wts.append(np.float32(scale * min(k, top)))
        ref = tab["pix"][rp[r]:rp[r + 1]]
        o = np.argsort(pix)
        np.testing.assert_array_equal(np.asarray(pix)[o], np.sort(ref))
        np.testing.assert_allclose(np.asarray(wts, np.float32)[o], tab["w"][rp[r]:rp[r + 1]][np.argsort(ref)],
                                   rtol=1e-6, atol=1e-7)
    grp, gcls = tab["sweep_grp"], tab["sweep_gcls"]
    assert grp.shape[1] * my <= 1024
    np.testing.assert_array_equal(np.sort(grp[grp >= 0]), idx)
    for row, c in zip(grp, gcls):
        assert np.all(cls[row[row >= 0]] == c)


@pytest.mark.parametrize("shape", [(14, 12, 17, (0, 360)), (16, 16, 17, (0, 180)), (99, 99, 103, (0, 180))])
def test_ray_pair_sweep_rule_reproduces_the_table(shape):
    """The ray-pair sweep's rule, emulated here with the kernel's arithmetic (float64 line geometry: the left ray's
    mid m1 L + mL0, p0 = ceil(mid - hw/|coef|) clamped to -2 .. len - 2; float32 weights clamp(hs - |ds (u + j)|, 0,
    ts) for the left ray's candidates p0 .. p0 + 2 and clamp(hs - |ds (u + j - 1/|coef|)|, 0, ts) for the right
    ray's p0 + 1 .. p0 + 3, u = p0 - mid), gives every ray of every angle the table's pixels and weights (pixels
    outside the image carry zero values in the kernel's padded lines); the pair groups hold each angle once, within
    512 threads."""
    nx, ny, my, ar = shape
    n_om = 16
    op = O.OracleOperator(cfg(nx=nx, ny=ny, my=my, n_omega=n_om, angle_range=ar, cor=0.7), (4.05,) * 3 + (90,) * 3,
                          "cubic", AL3, 0.35, orientations(2, 1), np.deg2rad(4.0), orthonormalize=False)
    idx = np.arange(n_om)
    g = _geo(op, "trapezoid")
    tab = radon_tables(g, idx, nx, ny, my, op.scale)
    rp, geo = tab["row_ptr"], tab["sweep_geo"]
    cls = (np.abs(g["X"][idx]) < np.abs(g["Y"][idx])).astype(int)
    f32 = np.float32
    for a in range(n_om):
        X, Y, T0, hw, top, scale = geo[a]
        coef, cross, ln, nl = (Y, X, ny, nx) if cls[a] else (X, Y, nx, ny)
        inv = 1.0 / coef
        ainv, hinv = abs(inv), hw * abs(inv)
        hs, ts, ds = f32(hw * scale), f32(top * scale), f32(coef * scale)
        c1, c2 = ds, f32(2) * ds
        cr = [f32((j - ainv) * ds) for j in (1.0, 2.0, 3.0)]
        got = np.zeros((my, nx * ny))
        for sp in range(-(-my // 2)):
            s = 2 * sp
            m1 = -cross * inv
            mL0 = -(T0 - s) * inv + min(inv, 0.0)
            bl, br = (s, s + 1) if inv > 0 else (s + 1, s)
            for L in range(nl):
                mid = m1 * L + mL0
                p0 = min(max(np.ceil(mid - hinv), -2.0), float(ln - 2))
                uu = f32(p0 - mid)
                wl = [np.clip(hs - abs(ds * uu), 0, ts), np.clip(hs - abs(f32(ds * uu + c1)), 0, ts),
                      np.clip(hs - abs(f32(ds * uu + c2)), 0, ts)]
                wr = [np.clip(hs - abs(f32(ds * uu + cr[j])), 0, ts) for j in range(3)]
                for j in range(4):
                    p = int(p0) + j
                    if not 0 <= p < ln:
                        continue
                    pix = p * nx + L if cls[a] else L * nx + p
                    if j <= 2 and bl < my:
                        got[bl, pix] += wl[j]
                    if j >= 1 and br < my:
                        got[br, pix] += wr[j - 1]
        ref = np.zeros((my, nx * ny))
        for s in range(my):
            r = a * my + s
            ref[s, tab["pix"][rp[r]:rp[r + 1]]] = tab["w"][rp[r]:rp[r + 1]]
        np.testing.assert_allclose(got, ref, rtol=1e-6, atol=1e-6 * float(ref.max()))
    grp = tab["sweep_grp_p"]
    assert grp[0, 0] >= 0 and grp.shape[1] * (-(-my // 2)) <= 512
    np.testing.assert_array_equal(np.sort(grp[grp >= 0]), idx)
    for row, c in zip(grp, tab["sweep_gcls_p"]):
        assert np.all(cls[row[row >= 0]] == c)
