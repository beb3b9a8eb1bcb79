"""FP64 model of the operator as a function of its continuous parameters, and its vector-Jacobian products (test
infrastructure: the checker of the product's PFGradient / ProjGradient and of set_pf_parameters /
set_proj_parameters; only tests/ import it). It follows the reference's differentiable model
(differentiable/reference.py:1-365, docs/differentiable_model.md), restated on top of oracle.model.OracleOperator:

    d[a, s, j]  = sum_k PF[a, k, j] S_k[a, s],         S_k[a, s] = scale sum_x K(t_a(x) - s) c_k(x)
    PF[a, k, j] = N(sigma_k) sum_nodes w sum_{axes h of ring p} m_h g(n_kh . v; sigma_k),   j = c P + p
    g(d; s)     = exp(-(1 - |d|) s) + exp(-(1 + |d|) s),  s = 1 / sigma^2, each term dropped at s(1 -+ |d|) >= cut
    n_kh        = exp([rot_k]x) U_k h                      (a left increment of orientation k)
    v           = Rot(omega)^T u(eta + eta_offset, theta_p),  u = cos(theta) (cos(eta) d0 + sin(eta) d90) - sin(theta) p0
    Rot(omega)  = exp(omega [k0']x),  k0' = exp([t1 e1 + t2 e2]x) k0   (axis tilt about e1, e2 perpendicular to k0)
    t_a(x)      = X_a x + Y_a y + T0_a - cor + shift_a       (detector coordinate in bins)

Parameters: rot (K, 3), sigma (K,), two_theta (P,), eta_offset, axis_tilt (2,), cor, shift (R,). The derivatives
are the analytic ones of the reference's model (reference.py:282-333): d g / d(n.v) = sign(d) s (e1 - e2),
d g / d s = -(1 - |d|) e1 - (1 + |d|) e2, d(n.v)/d rot = n x v, d v / d tilt_i = e_i x v - Rot^T (e_i x u),
d K / d t of the projector footprint. Pinned by tests/test_oracle_diffmodel.py: forward == OracleOperator.direct at
the initial parameters, and every group's vjp == central finite differences of forward (no cut, so the model is
smooth), including parameters away from their initial values.
"""
import numpy as np
from scipy.spatial.transform import Rotation

GROUPS = ("c", "rot", "sigma", "two_theta", "eta_offset", "axis_tilt", "cor", "shift")
PF_GROUPS = ("rot", "sigma", "eta_offset", "axis_tilt", "two_theta")


def perp_basis(k):
    """Two unit vectors e1, e2 perpendicular to k (and to each other) (reference.py:34-39: the same choice, so that
    axis_tilt means the same rotation)."""
    k = np.asarray(k, np.float64)
    k = k / np.linalg.norm(k)
    a = np.array([1.0, 0.0, 0.0]) if abs(k[0]) < 0.9 else np.array([0.0, 1.0, 0.0])
    e1 = np.cross(k, a)
    e1 /= np.linalg.norm(e1)
    return e1, np.cross(k, e1)


class OracleDiffModel:
    """The model of an OracleOperator as a function of its parameters. cut: "operator" (the operator's cut) or None
    (no cut: smooth, for finite differences)."""

    def __init__(self, orc, cut="operator"):
        self.o = orc
        self.cut = orc.cut if cut == "operator" else cut
        cfg = orc.cfg
        vec = lambda key: np.asarray(cfg[key], np.float64)
        self.k0, self.p0 = vec("k_direction_0"), vec("p_direction_0")
        self.d0, self.d90 = vec("detector_direction_origin"), vec("detector_direction_positive_90")
        self.e_tilt = perp_basis(self.k0)
        self.om = orc.om_nodes.reshape(orc.R, orc.S)
        self.eta = orc.eta_nodes.reshape(orc.C, orc.Se)
        yy, xx = np.meshgrid(np.arange(orc.Ny), np.arange(orc.Nx), indexing="ij")
        self.px, self.py = xx.ravel().astype(np.float64), yy.ravel().astype(np.float64)

    def params0(self):
        o = self.o
        return {"rot": np.zeros((o.K, 3)), "sigma": o.sigma.copy(), "two_theta": o.two_theta.astype(np.float64),
                "eta_offset": 0.0, "axis_tilt": np.zeros(2), "cor": o.cor, "shift": np.zeros(o.R)}

    # ------------------------------------------------------------------ geometry
    def rotations(self, prm):
        """Rot (R, S, 3, 3) of every omega node about the tilted axis."""
        t = np.asarray(prm["axis_tilt"], np.float64)
        k = Rotation.from_rotvec(t[0] * self.e_tilt[0] + t[1] * self.e_tilt[1]).as_matrix() @ self.k0
        o = self.o
        return Rotation.from_rotvec(self.om.reshape(-1, 1) * k).as_matrix().reshape(o.R, o.S, 3, 3)

    def u(self, prm):
        """u (C, Se, P, 3) and its derivatives with respect to eta_offset and to each ring's 2 theta."""
        eta = self.eta + float(prm["eta_offset"])
        e = np.cos(eta)[..., None] * self.d0 + np.sin(eta)[..., None] * self.d90
        de = -np.sin(eta)[..., None] * self.d0 + np.cos(eta)[..., None] * self.d90
        th = 0.5 * np.asarray(prm["two_theta"], np.float64)
        ct, st = np.cos(th)[None, None, :, None], np.sin(th)[None, None, :, None]
        u = ct * e[:, :, None, :] - st * self.p0
        return u, ct * de[:, :, None, :], 0.5 * (-st * e[:, :, None, :] - ct * self.p0)

    def probed(self, prm):
        u, _, _ = self.u(prm)
        return np.einsum("rsij,cepi->rscepj", self.rotations(prm), u)

    def probed_derivatives(self, prm):
        """d v / d eta_offset, d v / d (2 theta_p) (ring p's slice), d v / d tilt_i, each shaped like probed()."""
        Rot = self.rotations(prm)
        u, du_eta, du_tth = self.u(prm)
        v = np.einsum("rsij,cepi->rscepj", Rot, u)
        out = {"eta_offset": np.einsum("rsij,cepi->rscepj", Rot, du_eta),
               "two_theta": np.einsum("rsij,cepi->rscepj", Rot, du_tth)}
        for i, e in enumerate(self.e_tilt):
            out[("axis_tilt", i)] = np.cross(e, v) - np.einsum("rsij,cepi->rscepj", Rot, np.cross(e, u))
        return out

    def poles(self, prm):
        rot = Rotation.from_rotvec(np.asarray(prm["rot"], np.float64)).as_matrix() @ self.o.U
        return np.einsum("kij,aj->kai", rot, self.o.axes)

    # ------------------------------------------------------------------ pole figure
    def g(self, d, s):
        """g, d g / d(n.v) and d g / d s for signed dot products d and s = 1 / sigma^2 (broadcast)."""
        ad = np.abs(d)
        x1, x2 = (1.0 - ad) * s, (1.0 + ad) * s
        e1, e2 = np.exp(-x1), np.exp(-x2)
        if self.cut is not None:
            e1 = np.where(x1 < self.cut, e1, 0.0)
            e2 = np.where(x2 < self.cut, e2, 0.0)
        return e1 + e2, np.sign(d) * s * (e1 - e2), -(1.0 - ad) * e1 - (1.0 + ad) * e2

    def pf(self, prm):
        """PF (R, K, C, P)."""
        o = self.o
        v, n = self.probed(prm), self.poles(prm)
        sig = np.asarray(prm["sigma"], np.float64)
        s = 1.0 / sig ** 2
        norm = 1.0 / (8.0 * np.pi * sig ** 2)
        out = np.zeros((o.R, o.K, o.C, o.P))
        for p in range(o.P):
            a0, a1 = o.start[p], o.start[p + 1]
            d = np.einsum("rscex,kax->rsceka", v[..., p, :], n[:, a0:a1])
            gv, _, _ = self.g(d, s[:, None])
            node = np.einsum("rsceka,a->rscek", gv, o.counts[a0:a1])
            out[..., p] = norm[None, :, None] * np.einsum("rscek,se->rkc", node, o.qw)
        return out

    # ------------------------------------------------------------------ projector
    def footprint(self, d, A):
        """K(d) and d K / d d of the operator's footprint (trapezoid, or hat)."""
        if self.o.kernel == "hat":
            return np.maximum(1.0 - np.abs(d), 0.0), np.where(np.abs(d) < 1.0, -np.sign(d), 0.0)
        hw, top = 0.5 * (1.0 + A), min(1.0, A)
        q = hw - np.abs(d)
        return np.clip(q, 0.0, top) / A, np.where((q > 0.0) & (q < top), -np.sign(d) / A, 0.0)

    def projection(self, prm):
        """P (R, My, Npix) and d P / d t: S_k[a] = P[a] @ c_k; d P / d t is the derivative for a shift of t_a."""
        o = self.o
        Pm = np.zeros((o.R, o.My, o.Nx * o.Ny))
        dPm = np.zeros_like(Pm)
        shift = np.asarray(prm["shift"], np.float64)
        T0 = o.T0 + o.cor                              # the operator's T0 without its centre of rotation
        for a in range(o.R):
            A = max(abs(o.X[a]), abs(o.Y[a]))
            t = o.X[a] * self.px + o.Y[a] * self.py + T0[a] - float(prm["cor"]) + shift[a]
            for s in range(o.My):
                k, dk = self.footprint(t - s, A)
                Pm[a, s] = o.scale * k
                dPm[a, s] = o.scale * dk
        return Pm, dPm

    # ------------------------------------------------------------------ model and derivatives
    def forward(self, c, prm):
        o = self.o
        PF = self.pf(prm).reshape(o.R, o.K, o.CP)
        Pm, _ = self.projection(prm)
        S = np.einsum("asp,kp->ask", Pm, np.asarray(c, np.float64).reshape(o.K, -1))
        return np.einsum("akj,ask->asj", PF, S)

    def vjp(self, c, prm, r, groups=GROUPS):
        """The gradients of <r, forward(c, prm)> with respect to the groups (c shaped as c, the others as prm)."""
        o = self.o
        r = np.asarray(r, np.float64).reshape(o.R, o.My, o.CP)
        cf = np.asarray(c, np.float64).reshape(o.K, -1)
        Pm, dPm = self.projection(prm)
        S = np.einsum("asp,kp->ask", Pm, cf)
        PF = self.pf(prm)
        Z = np.einsum("akj,asj->ask", PF.reshape(o.R, o.K, o.CP), r)          # PF^T r per angle
        out = {}
        if "c" in groups:
            out["c"] = np.einsum("asp,ask->kp", Pm, Z).reshape(np.shape(c))
        if "shift" in groups or "cor" in groups:
            gs = np.einsum("asp,kp,ask->a", dPm, cf, Z)
            if "shift" in groups:
                out["shift"] = gs
            if "cor" in groups:
                out["cor"] = -float(gs.sum())
        pg = [g_ for g_ in groups if g_ in PF_GROUPS]
        if pg:
            G = np.einsum("asj,ask->akj", r, S).reshape(o.R, o.K, o.C, o.P)     # dL / dPF
            out.update(self.pf_grads(prm, G, pg))
        return out

    def pf_grads(self, prm, G, groups, curv=None, S=None):
        """sum over (a, k, c, p) of G dPF / d(parameter), for the PF groups."""
        o = self.o
        v, n = self.probed(prm), self.poles(prm)
        dv = self.probed_derivatives(prm)
        sig = np.asarray(prm["sigma"], np.float64)
        s = 1.0 / sig ** 2
        norm = 1.0 / (8.0 * np.pi * sig ** 2)
        out = {"rot": np.zeros((o.K, 3)), "sigma": np.zeros(o.K), "two_theta": np.zeros(o.P),
               "eta_offset": 0.0, "axis_tilt": np.zeros(2)}
        for p in range(o.P):
            a0, a1 = o.start[p], o.start[p + 1]
            vp, npk = v[..., p, :], n[:, a0:a1]
            d = np.einsum("rscex,kax->rsceka", vp, npk)
            gv, dgd, dgs = self.g(d, s[:, None])
            m = o.counts[a0:a1]
            # the weight of every node term: G * N(sigma_k) * w_node, (R, S, C, Se, K)
            W = G[..., p].transpose(0, 2, 1)[:, None, :, None, :] * o.qw[None, :, None, :, None] * norm
            Wd = W[..., None] * m * dgd                                           # (R, S, C, Se, K, A_p)
            out["rot"] += np.einsum("rsceka,rscekax->kx", Wd,
                                    np.cross(npk[None, None, None, None], vp[:, :, :, :, None, None, :]))
            node = np.einsum("rsceka,a->rscek", gv, m)
            dnode = np.einsum("rsceka,a->rscek", dgs, m)
            out["sigma"] += np.einsum("rscek,rscek->k", W, -2.0 / sig * node - 2.0 / sig ** 3 * dnode)
            proj = lambda arr: np.einsum("rsceka,rsceka->", Wd, np.einsum("rscex,kax->rsceka", arr[..., p, :], npk))
            out["eta_offset"] += proj(dv["eta_offset"])
            out["two_theta"][p] += proj(dv["two_theta"])
            for i in range(2):
                out["axis_tilt"][i] += proj(dv[("axis_tilt", i)])
        return {k: out[k] for k in groups}

    def pf_jacobian_blocks(self, prm, curv, S):
        """The Gauss-Newton blocks H_k = J_k^T diag(curv) J_k (K, 4, 4) for the rotation (3) and width of every
        orientation: J_k[a, s, j] = S[a, s, k] dPF[a, k, j] / d(rot_k, sigma_k)."""
        o = self.o
        v, n = self.probed(prm), self.poles(prm)
        sig = np.asarray(prm["sigma"], np.float64)
        s = 1.0 / sig ** 2
        norm = 1.0 / (8.0 * np.pi * sig ** 2)
        C2 = np.einsum("asj,ask->akj", np.asarray(curv, np.float64).reshape(o.R, o.My, o.CP), S * S)
        H = np.zeros((o.K, 4, 4))
        for p in range(o.P):
            a0, a1 = o.start[p], o.start[p + 1]
            vp, npk = v[..., p, :], n[:, a0:a1]
            d = np.einsum("rscex,kax->rsceka", vp, npk)
            gv, dgd, dgs = self.g(d, s[:, None])
            m = o.counts[a0:a1]
            w = o.qw[None, :, None, :, None] * norm                               # (1, S, 1, Se, K)
            J = np.zeros((o.R, o.C, o.K, 4))
            cr = np.cross(npk[None, None, None, None], vp[:, :, :, :, None, None, :])
            J[..., :3] = np.einsum("rsceka,rscekax->rckx", (w[..., None] * m * dgd), cr)
            node = np.einsum("rsceka,a->rscek", gv, m)
            dnode = np.einsum("rsceka,a->rscek", dgs, m)
            J[..., 3] = np.einsum("rscek->rck", w * (-2.0 / sig * node - 2.0 / sig ** 3 * dnode))
            cp = C2.reshape(o.R, o.K, o.C, o.P)[..., p].transpose(0, 2, 1)        # (R, C, K)
            H += np.einsum("rck,rckx,rcky->kxy", cp, J, J)
        return H


def perturb(prm, group, index, h):
    """A copy of prm with prm[group][index] (or the scalar prm[group]) moved by h."""
    out = {k: (np.array(v, dtype=np.float64, copy=True) if np.ndim(v) else float(v)) for k, v in prm.items()}
    if np.ndim(out[group]) == 0:
        out[group] = out[group] + h
    else:
        out[group][index] += h
    return out
