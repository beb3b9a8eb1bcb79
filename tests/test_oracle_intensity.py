"""The oracle's intensity layer (differentiable/data_model.py:20-64) and smoothed TV, pinned by what the
mathematics fixes: one-hot scales touch exactly their rows, points or segments (a wrong broadcast or ring index
fails), the sum over unit inputs has a closed form, and the vjp equals the exact directional derivative of the
(multi-linear) layer."""
import numpy as np

from oracle.model import intensity_forward, intensity_vjp, smooth_tv

R, MY, C, P = 5, 4, 3, 2


def _prm(rng):
    return (rng.uniform(0.5, 1.5, R), rng.uniform(0.5, 1.5, MY), rng.uniform(0.5, 1.5, P), rng.uniform(0, 1, C * P))


def test_one_hot_scales_touch_their_own_entries():
    u = np.ones((R, MY, C * P))
    z = np.zeros
    for a in range(R):
        f = z(R)
        f[a] = 2.0
        d = intensity_forward(u, f, np.ones(MY), np.ones(P), z(C * P), P)
        assert np.all(d[a] == 2.0) and np.all(np.delete(d, a, axis=0) == 0.0)
    for s_ in range(MY):
        t = z(MY)
        t[s_] = 3.0
        d = intensity_forward(u, np.ones(R), t, np.ones(P), z(C * P), P)
        assert np.all(d[:, s_] == 3.0) and np.all(np.delete(d, s_, axis=1) == 0.0)
    for p in range(P):
        rp = z(P)
        rp[p] = 5.0
        d = intensity_forward(u, np.ones(R), np.ones(MY), rp, z(C * P), P)
        seg = np.arange(C * P)
        assert np.all(d[:, :, seg % P == p] == 5.0) and np.all(d[:, :, seg % P != p] == 0.0)


def test_sum_closed_form():
    rng = np.random.default_rng(0)
    f, t, rp, bg = _prm(rng)
    d = intensity_forward(np.ones((R, MY, C * P)), f, t, rp, bg, P)
    assert np.isclose(d.sum(), f.sum() * t.sum() * C * rp.sum() + R * MY * bg.sum(), rtol=1e-13)


def test_vjp_is_the_directional_derivative():
    rng = np.random.default_rng(1)
    f, t, rp, bg = _prm(rng)
    u = rng.random((R, MY, C * P))
    r = rng.standard_normal((R, MY, C * P))
    g = intensity_vjp(u, f, t, rp, r, P)
    L = lambda f_, t_, rp_, bg_, u_: float((intensity_forward(u_, f_, t_, rp_, bg_, P) * r).sum())
    # the layer is linear in each group: the directional derivative is exactly L(x + v) - L(x) per group
    dirs = {"angle_scale": rng.standard_normal(R), "trans_scale": rng.standard_normal(MY),
            "ring_scale": rng.standard_normal(P), "background": rng.standard_normal(C * P),
            "u": rng.standard_normal(u.shape)}
    base = L(f, t, rp, bg, u)
    args = {"angle_scale": f, "trans_scale": t, "ring_scale": rp, "background": bg, "u": u}
    for k, v in dirs.items():
        a2 = dict(args)
        a2[k] = args[k] + v
        fd = L(a2["angle_scale"], a2["trans_scale"], a2["ring_scale"], a2["background"], a2["u"]) - base
        assert np.isclose(fd, float((g[k] * v).sum()), rtol=1e-10, atol=1e-10), k


def test_smooth_tv_closed_forms():
    c = np.full((2, 4, 5), 0.7)
    assert np.isclose(smooth_tv(c, eps=1e-8), 2 * 4 * 5 * 1e-4)             # constant: sqrt(eps) per pixel
    ramp = np.broadcast_to(np.arange(5.0) * 0.5, (1, 4, 5)).copy()           # x ramp, slope 0.5
    # pixels with a right neighbour: |dx| = 0.5, the last column: sqrt(eps)
    assert np.isclose(smooth_tv(ramp, eps=0.0), 4 * 4 * 0.5)
    diag = (np.arange(4.0)[:, None] + np.arange(5.0)[None, :])[None]         # dx = dy = 1 inside
    assert np.isclose(smooth_tv(diag, eps=0.0), 3 * 4 * np.sqrt(2) + 3 + 4)
