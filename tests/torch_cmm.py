"""Plain-torch restatement of the barotropic CMM loop, for an autograd
gradient that is independent of the hand-written adjoint kernels.

Test infrastructure only.  The differentiable arithmetic (Powell-Sabin
assembly, split-triangle evaluation, the RK4 / Euler / improved-Euler feet,
the stencil projection to jets, composition and vorticity sampling) is
written here in torch ops from the oracle's NumPy formulas
(oracle/splines.py, oracle/evolution.py).  The linear spectral pipelines are
dense matrices obtained by applying the oracle to unit vectors:

* grid vorticity -> velocity jets on the MWSS mesh: oracle inv_Laplacian +
  project_stream (spherical_harmonic_tools.py:14-19, 66-105),
* coefficients -> initial-vorticity jets: oracle project_onto_S12
  (spherical_harmonic_tools.py:108-142),
* grid -> [l, m >= 0] coefficients: oracle forward_lm.

Point location (a discrete choice) comes from the oracle's locators on
detached positions; autograd then differentiates the loop for those
locations, as the adjoint does.  The loop order (bootstrap, stepping, remaps,
time levels) is spherical_simulations.py:216-294.
"""
import numpy as np
import torch

from oracle import meshes as om, sht as osht, splines as osp

F64 = torch.float64
# Cs table (spherical_spline.py:660-662)
_CS = osp.Cs


def _t(x):
    return torch.as_tensor(np.asarray(x), dtype=F64)


def assemble(vals, simplices, ps):
    """oracle.splines.assemble in torch: vals (C,4,Nv) or (4,Nv) -> (C,19,T) / (19,T)."""
    Hs, Gs, Es = [[_t(h) for h in x] for x in ps]
    simp = torch.as_tensor(np.asarray(simplices), dtype=torch.long)
    scalar = vals.dim() == 2
    if scalar:
        tv = vals[:, simp]                          # (4, T, 3)
        f = [tv[0, :, i] for i in range(3)]

        def hdot(h, i):                             # h (3, T) . grad (3, T)
            return (h * tv[1:, :, i]).sum(0)
    else:
        tv = vals[:, :, simp]                       # (C, 4, T, 3)
        f = [tv[:, 0, :, i] for i in range(3)]

        def hdot(h, i):                             # sum_d h_d vals[d, 1+c]
            return torch.einsum("dt,dct->ct", h, tv[:, 1:, :, i])
    c = [None] * 19

    def vert(G, col, H, i):
        return (1 / G[:, col]) * (hdot(H, i) / 2 - G[:, 0] * f[i])

    c[0], c[1], c[2] = f[0], f[1], f[2]
    c[3] = vert(Gs[0], 1, Hs[0], 0)
    c[4] = vert(Gs[6], 1, Hs[6], 0)
    c[5] = vert(Gs[4], 2, Hs[4], 0)
    c[6] = vert(Gs[2], 1, Hs[2], 1)
    c[7] = vert(Gs[7], 1, Hs[7], 1)
    c[8] = vert(Gs[1], 2, Hs[1], 1)
    c[9] = vert(Gs[5], 1, Hs[5], 2)
    c[10] = vert(Gs[8], 1, Hs[8], 2)
    c[11] = vert(Gs[3], 2, Hs[3], 2)
    c[12] = Es[0][:, 0] * c[3] + Es[0][:, 1] * c[8]
    c[13] = Es[1][:, 1] * c[6] + Es[1][:, 2] * c[11]
    c[14] = Es[2][:, 2] * c[9] + Es[2][:, 0] * c[5]
    c[15] = Es[0][:, 0] * c[4] + Es[0][:, 1] * c[7]
    c[16] = Es[1][:, 1] * c[7] + Es[1][:, 2] * c[10]
    c[17] = Es[2][:, 0] * c[4] + Es[2][:, 2] * c[10]
    c[18] = Gs[9][:, 0] * c[4] + Gs[9][:, 1] * c[7] + Gs[9][:, 2] * c[10]
    return torch.stack(c, -2)


def _det(a, b, c):
    return (a * torch.cross(b, c, dim=-1)).sum(-1)


def _bary(v1, v2, v3, q):
    d = _det(v1, v2, v3)
    return torch.stack([_det(q, v2, v3) / d, _det(v1, q, v3) / d, _det(v1, v2, q) / d], -1)


class Located:
    """A located point set: triangle, split pick and split-triangle vertices
    (constants) from the detached positions."""

    def __init__(self, tri, vs, q_np):
        self.tri = torch.as_tensor(tri, dtype=torch.long)
        b = np.asarray(osp_bary(vs, q_np))
        mn, mx = np.argmin(b, axis=1), np.argmax(b, axis=1)
        self.cs = torch.as_tensor(_CS[mn, mx], dtype=torch.long)           # (N, 6)
        self.vn = _t(osp.new_vpoints(vs, (mn, mx)))                        # (N, 3, 3)

    def eval(self, coeffs, q):
        """un-normalised PS value(s) at q (N,3): coeffs (..., 19, T) -> (..., N)."""
        bn = _bary(self.vn[:, 0], self.vn[:, 1], self.vn[:, 2], q)
        cf = coeffs[..., self.tri]                                          # (..., 19, N)
        idx = self.cs.t()                                                   # (6, N)
        sel = lambda j: torch.gather(cf, -2, idx[j].expand(cf.shape[:-2] + (1, -1)).contiguous()).squeeze(-2)
        return ((bn[:, 0] ** 2) * sel(0) + (bn[:, 1] ** 2) * sel(1) + (bn[:, 2] ** 2) * sel(2)
                + 2 * (bn[:, 1] * bn[:, 0]) * sel(3) + 2 * (bn[:, 2] * bn[:, 1]) * sel(4)
                + 2 * (bn[:, 2] * bn[:, 0]) * sel(5))


def osp_bary(vs, q):
    from oracle import geometry as g
    return g.bary_coords(vs[:, 0], vs[:, 1], vs[:, 2], q)


class TorchCMM:
    """J(p) of the loop of MisfitProblem, in torch (CPU, FP64)."""

    EPS = 1e-5

    def __init__(self, L, Nt, T, vertices, simplices, target_lm, rot_rate=2 * np.pi, n_maps=10, lmax=4):
        self.L, self.Nt, self.T, self.rot, self.lmax = L, Nt, T, float(rot_rate), lmax
        self.V = np.asarray(vertices, dtype=np.float64)
        self.S = np.asarray(simplices, dtype=np.int64)
        self.ps = om.ps_split(self.V, self.S)
        s_pts, tan = om.stencil_points(self.V, self.EPS)
        self.spts = _t(s_pts)                                              # (3, 4, Nv)
        self.a1, self.a2 = _t(tan[0]), _t(tan[1])                           # (Nv, 3)
        self.mu = om.StructuredTriangulation(L)
        self.samples_np = self.mu.sample_points()
        self.samples = _t(self.samples_np)
        self.cos_grid = _t(np.repeat(np.cos(om.mwss_thetas(L)), 2 * L))
        self.ns = [i for i in range(1, Nt) if i % n_maps == 0] + [Nt]
        self.tspan = np.linspace(0, T, Nt, endpoint=False)
        self.dt = self.tspan[1] - self.tspan[0]
        self.target = target_lm                                             # (L, L) complex [l, m]
        cm = np.full((L, L), 2.0)
        cm[:, 0] = 1.0
        self.cm = _t(cm)
        self._build_matrices()

    # ---- dense linear maps from the oracle ---------------------------------
    def _build_matrices(self):
        L = self.L
        G = (L + 1) * 2 * L
        D_inv, Lp, Lm, Lz = osht.operator_tables(L)
        Nu = len(self.mu.vertices)
        Mv = np.zeros((3, 4, Nu, G))
        temp = np.zeros((L, 2 * L + 1), dtype=np.complex128)
        for k in range(G):
            e = np.zeros(G)
            e[k] = 1.0
            psi = osht.inv_Laplacian(e.reshape(L + 1, 2 * L), L, D_inv)
            U = osp.VelocitySpline(self.mu, np.zeros((3, 3, 4, Nu)), np.zeros((3, 3, 19, len(self.mu.simplices))),
                                   ts=self.tspan[0:3])
            osht.project_stream(psi, L, Lp, Lm, Lz, U, 0, temp)
            Mv[..., k] = U.vals[0]
        self.Mv = _t(Mv)
        # coefficients (1 <= l <= lmax, m >= 0; Re, then Im for m > 0) -> w0 jets
        cols = []
        for l in range(1, self.lmax + 1):
            for m in range(0, l + 1):
                for part in ((1.0,) if m == 0 else (1.0, 1j)):
                    a = np.zeros((L, L), dtype=np.complex128)
                    a[l, m] = part
                    sp = osht.project_onto_S12(osht.lm_to_full(a, L), L, self.mu)
                    cols.append(sp.vals.copy())
        self.Mw = _t(np.stack(cols, -1))                                   # (4, Nu, P)
        # analysis: grid -> [l, m] complex (re, im)
        A = np.zeros((2, L, L, G))
        for k in range(G):
            e = np.zeros(G)
            e[k] = 1.0
            f = osht.forward_lm(e.reshape(L + 1, 2 * L), L)
            A[0, :, :, k], A[1, :, :, k] = f.real, f.imag
        self.A = _t(A)

    # ---- pieces -------------------------------------------------------------
    def _loc_ico(self, q):
        qn = q.detach().numpy()
        tri, _ = om.locate_kdtree(self.V, self.S, qn)
        return Located(tri, self.V[self.S[tri]], qn)

    def _loc_mwss(self, q):
        qn = q.detach().numpy()
        _, tri, vs = self.mu.query(qn)
        return Located(tri, vs, qn)

    def _map_eval(self, coeffs, q):
        r = self._loc_ico(q).eval(coeffs, q)                               # (3, N)
        return (r / r.norm(dim=0)).t()

    def _compose(self, chain, q):
        x = q
        for c in chain:
            x = self._map_eval(c, x)
        return x

    def _sample(self, x, rot):
        nx = x.norm(dim=1, keepdim=True)
        u = x / nx
        w0 = self._loc_mwss(u).eval(self.w0c, u)
        return w0 + 2 * rot * u[:, 2] - 2 * rot * self.cos_grid

    def _velocity(self, X, rot):
        w = self._sample(X, rot)
        vals = torch.einsum("cknj,j->ckn", self.Mv, w)
        return assemble(vals, self.mu.simplices, self.mu.ps_split)         # (3, 19, Tu)

    def _vel_eval(self, levels, nVs, ts0, t, x):
        loc = self._loc_mwss(x)
        w = osp.lagrange_weights(nVs, t, self.dt, ts0) if nVs > 1 else (1.0, 0.0, 0.0)
        out = 0
        for k in range(min(nVs, 3)):
            if w[k] != 0.0:
                out = out + w[k] * loc.eval(levels[k], x)
        return out.t()                                                      # (N, 3)

    def _feet(self, kind, t, levels, nVs, ts0):
        x = self.spts.permute(1, 2, 0).reshape(-1, 3)                       # (4 Nv, 3) stencil points
        nrm = lambda z: z / z.norm(dim=1, keepdim=True)
        dt = self.dt
        V = lambda tt, y: self._vel_eval(levels, nVs, ts0, tt, y)
        if kind == "euler":
            return nrm(x - dt * V(t, x))
        if kind == "improved_euler":
            v1 = V(t, x)
            s1 = nrm(x - dt * v1)
            v2 = V(t + dt, s1)
            return nrm(x - dt / 2 * (v1 + v2))
        v1 = V(t + dt, x)
        s1 = nrm(x - dt * v1 / 2)
        v2 = V(t + dt / 2, s1)
        s2 = nrm(x - dt / 2 * v2)
        v3 = V(t + 0.5 * dt, s2)
        s3 = nrm(x - dt * v3)
        v4 = V(t, s3)
        return nrm(x - dt / 6 * (v1 + 2 * v2 + 2 * v3 + v4))

    def _adv(self, cur, kind, t, levels, nVs, ts0):
        """map update (advect + spline_proj_sphere, evolution_functions.py:90-142)."""
        Nv = len(self.V)
        ys = self._feet(kind, t, levels, nVs, ts0).reshape(4, Nv, 3)        # [k, v, c]
        if cur is not None:
            # vertex feet locate the split triangle the four stencil feet use
            vv = _t(self.V)
            yv = self._vertex_feet(kind, t, levels, nVs, ts0, vv)
            loc = self._loc_ico(yv)
            S = torch.stack([loc.eval(cur, ys[k]).t() for k in range(4)])   # (4, Nv, 3), not normalised
        else:
            S = ys
        mean = S.mean(0)
        val = mean / mean.norm(dim=1, keepdim=True)
        e4 = 4 * self.EPS
        d1 = (S[1] - S[0] + S[3] - S[2]) / e4
        d2 = (S[2] - S[0] + S[3] - S[1]) / e4
        # jets[c][0] = value_c, jets[d][1 + c] = d1_c a1_d + d2_c a2_d (J^T layout)
        jets = torch.cat([val.t()[:, None, :],
                          torch.einsum("nc,nd->dcn", d1, self.a1) + torch.einsum("nc,nd->dcn", d2, self.a2)], 1)
        return assemble(jets, self.S, self.ps)                              # (3, 19, T)

    def _vertex_feet(self, kind, t, levels, nVs, ts0, vv):
        """vertex feet (RK4, the only integrator of a non-identity update);
        they only pick the split triangle, so no gradient flows through them"""
        nrm = lambda z: z / z.norm(dim=1, keepdim=True)
        dt = self.dt
        x = vv.detach()
        V = lambda tt, y: self._vel_eval(levels, nVs, ts0, tt, y)
        with torch.no_grad():
            v1 = V(t + dt, x)
            s1 = nrm(x - dt * v1 / 2)
            v2 = V(t + dt / 2, s1)
            s2 = nrm(x - dt / 2 * v2)
            v3 = V(t + 0.5 * dt, s2)
            s3 = nrm(x - dt * v3)
            v4 = V(t, s3)
            return nrm(x - dt / 6 * (v1 + 2 * v2 + 2 * v3 + v4))

    # ---- the loop -------------------------------------------------------------
    def J(self, p):
        """p: (P,) float64 tensor (requires_grad for the gradient)."""
        L = self.L
        vals_w = torch.einsum("knp,p->kn", self.Mw, p)
        self.w0c = assemble(vals_w, self.mu.simplices, self.mu.ps_split)   # (19, Tu)
        dt = self.dt
        q = self.samples
        U0 = self._velocity(q, 0.0)
        M0 = self._adv(None, "euler", 0.0, [U0], 1, 0.0)
        U1 = self._velocity(self._compose([M0], q), self.rot)
        M1 = self._adv(None, "improved_euler", 0.0, [U0, U1], 2, 0.0)
        U2 = self._velocity(self._compose([M1], q), self.rot)
        M2 = self._adv(M1, "rk4", dt, [U0, U2], 2, 0.0)
        U3 = self._velocity(self._compose([M2], q), self.rot)
        levels, nVs, ts0 = [U0, U2, U3], 3, 0.0
        curr, submaps, count = M2, [], 2
        while count < self.Nt:
            t = self.tspan[count]
            curr = self._adv(curr, "rk4", t, levels, nVs, ts0)
            Un = self._velocity(self._compose([curr] + submaps[::-1], q), self.rot)
            levels = levels[1:] + [Un]
            ts0 = float(np.float64(ts0) + dt)
            count += 1
            if count in self.ns:
                submaps.append(curr)
                curr = None
        chain = ([] if curr is None else [curr]) + submaps[::-1]
        w = self._sample(self._compose(chain, q), self.rot)
        coef = torch.einsum("rlmj,j->rlm", self.A, w)
        d_re = coef[0] - _t(self.target.real)
        d_im = coef[1] - _t(self.target.imag)
        mask = _t(np.tril(np.ones((L, L))))                                 # m <= l
        return (self.cm * mask * (d_re ** 2 + d_im ** 2)).sum()
