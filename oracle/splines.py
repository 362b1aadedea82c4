"""Powell-Sabin split splines S^1_2(T_PS) (oracle).
Restates cmm_s2/core/spherical_spline.py.

Pinned by tests/test_oracle_splines.py:
* exact reproduction of homogeneous quadratics p(x) = x^T A x on any mesh
  (raw, un-normalised evaluation equals p(q) to round-off, also when
  extrapolating as stencil_eval does) -- this pins every index of ``Cs`` /
  ``edges_ps`` and every coefficient formula of ``assemble_coefficients``;
* vertex interpolation and C^0/C^1 continuity across split and macro edges;
* O(h^3) convergence on f = x^5 + xyz + sin y
  (tests/test_spline_interpolant.py:40-47 of the reference);
* map normalisation |X| = 1; Lagrange weights reproduce quadratics in t.
"""
import numpy as np

from . import geometry as g

# spherical_spline.py:660-666
Cs = np.array([[[0, 0, 0, 0, 0, 0], [18, 1, 13, 7, 6, 16], [13, 2, 18, 11, 10, 16]],
               [[18, 14, 0, 17, 5, 4], [0, 0, 0, 0, 0, 0], [14, 18, 2, 17, 10, 9]],
               [[0, 12, 18, 3, 15, 4], [1, 18, 12, 7, 15, 8], [0, 0, 0, 0, 0, 0]]])

edges_ps = np.array([[[0, 0, 0], [3, 1, 5], [5, 2, 3]],
                     [[3, 6, 0], [0, 0, 0], [6, 3, 2]],
                     [[0, 4, 3], [1, 3, 4], [0, 0, 0]]])


def bary_minmax(X):
    """spherical_spline.py:691-697 (first occurrence on ties)."""
    return np.argmin(X, axis=1), np.argmax(X, axis=1)


def new_vpoints(v_pts, bmm):
    """spherical_spline.py:668-689 -- vertices of the split sub-triangle."""
    v1, v2, v3 = v_pts[:, 0, :].copy(), v_pts[:, 1, :].copy(), v_pts[:, 2, :].copy()
    e1 = g.div_norm(v1 / 2 + v2 / 2).T
    e2 = g.div_norm(v2 / 2 + v3 / 2).T
    e3 = g.div_norm(v3 / 2 + v1 / 2).T
    v4 = g.div_norm((v1 + v2 + v3) / 3).T
    EE = np.array([v1, v2, v3, v4, e1, e2, e3])
    v_pts_n = np.empty(np.shape(v_pts))
    Js = edges_ps[bmm[0], bmm[1]]
    N = np.shape(Js)[0]
    r = np.arange(N)
    v_pts_n[:, 0, :] = EE[Js[:, 0], r, :]
    v_pts_n[:, 1, :] = EE[Js[:, 1], r, :]
    v_pts_n[:, 2, :] = EE[Js[:, 2], r, :]
    return v_pts_n


def assemble(vals, simplices, ps):
    """spherical_spline.py:300-337 (same formulas at 115-152, 471-508, 619-655).

    vals: (..., 4, Nv) jets -- vals[..., 0, v] = f(v), and for the map/velocity
    the leading axis is the component with J^T storage vals[d, 1+c] = d_d f_c.
    The contraction ``utils.dot`` over axis 0 of ``tri_vals[:, 1::]`` is what
    turns that storage into h . grad f_c.  Returns coeffs (..., 19, Ntri).
    """
    Hs, Gs, Es = ps
    simplices = np.array(simplices)
    vals = np.asarray(vals)
    scalar = vals.ndim == 2
    if scalar:
        tv = vals[:, simplices]                     # (4, T, 3)
        f = [tv[0, :, i] for i in range(3)]

        def hdot(h, i):
            return g.dot(h, tv[1:, :, i])
    else:
        tv = vals[:, :, simplices]                  # (3, 4, T, 3)
        f = [tv[:, 0, :, i] for i in range(3)]

        def hdot(h, i):
            return g.dot(h, tv[:, 1:, :, i])
    shape = f[0].shape
    c = np.zeros(shape[:-1] + (19,) + shape[-1:])

    def put(k, x):
        c[..., k, :] = x

    def vert(k, G, col, H, i):
        put(k, (1 / G[:, col]) * (hdot(H, i) / 2 - G[:, 0] * f[i]))

    put(0, f[0]); put(1, f[1]); put(2, f[2])
    vert(3, Gs[0], 1, Hs[0], 0)
    vert(4, Gs[6], 1, Hs[6], 0)
    vert(5, Gs[4], 2, Hs[4], 0)
    vert(6, Gs[2], 1, Hs[2], 1)
    vert(7, Gs[7], 1, Hs[7], 1)
    vert(8, Gs[1], 2, Hs[1], 1)
    vert(9, Gs[5], 1, Hs[5], 2)
    vert(10, Gs[8], 1, Hs[8], 2)
    vert(11, Gs[3], 2, Hs[3], 2)
    C = lambda k: c[..., k, :]
    put(12, Es[0][:, 0] * C(3) + Es[0][:, 1] * C(8))
    put(13, Es[1][:, 1] * C(6) + Es[1][:, 2] * C(11))
    put(14, Es[2][:, 2] * C(9) + Es[2][:, 0] * C(5))
    put(15, Es[0][:, 0] * C(4) + Es[0][:, 1] * C(7))
    put(16, Es[1][:, 1] * C(7) + Es[1][:, 2] * C(10))
    put(17, Es[2][:, 0] * C(4) + Es[2][:, 2] * C(10))
    put(18, Gs[9][:, 0] * C(4) + Gs[9][:, 1] * C(7) + Gs[9][:, 2] * C(10))
    return c


def _bb(bcc_n, cfs, nCs):
    """The 6-term quadratic Bernstein-Bezier sum (spherical_spline.py:196-201).
    cfs: (..., 19, N) coefficients already gathered at the located triangles."""
    r = np.arange(len(nCs))
    sel = lambda j: cfs[..., nCs[:, j], r]
    return ((bcc_n[:, 0] ** 2) * sel(0) + (bcc_n[:, 1] ** 2) * sel(1)
            + (bcc_n[:, 2] ** 2) * sel(2) + 2 * (bcc_n[:, 1] * bcc_n[:, 0]) * sel(3)
            + 2 * (bcc_n[:, 2] * bcc_n[:, 1]) * sel(4) + 2 * (bcc_n[:, 2] * bcc_n[:, 0]) * sel(5))


def located_eval(query, coeffs, q_pts):
    """Query + split pick + BB evaluation, un-normalised.
    query: mesh.query callable; coeffs (..., 19, T); q_pts (N,3)."""
    bcc, trangs, v_pts = query(q_pts)
    bb = bary_minmax(bcc)
    nCs = Cs[bb[0], bb[1]]
    v_pts_n = new_vpoints(v_pts, bb)
    cfs = coeffs[..., trangs]
    bcc_n = g.bary_coords(v_pts_n[:, 0, :], v_pts_n[:, 1, :], v_pts_n[:, 2, :], q_pts)
    return _bb(bcc_n, cfs, nCs)


def map_eval(query, coeffs, q_pts):
    """sphere_diffeomorphism.__call__ (spherical_spline.py:178-205) -> (3,N)."""
    outs = located_eval(query, coeffs, q_pts)
    norm = np.sqrt(outs[0, :] ** 2 + outs[1, :] ** 2 + outs[2, :] ** 2)
    return (1 / norm)[None, :] * outs


def stencil_eval(query, coeffs, q_pts, st_pts):
    """sphere_diffeomorphism.stencil_eval (spherical_spline.py:268-298).
    q_pts (3,Nv) vertex feet; st_pts (3,4,Nv) stencil feet, overwritten."""
    bcc, trangs, v_pts = query(q_pts.T)
    bmm = bary_minmax(bcc)
    v_pts_n = new_vpoints(v_pts, bmm)
    cfs = coeffs[:, :, trangs]
    nCs = Cs[bmm[0], bmm[1]]
    st_pts_n = st_pts[:]
    for i in range(4):
        bcc_n = g.bary_coords(v_pts_n[:, 0, :], v_pts_n[:, 1, :], v_pts_n[:, 2, :],
                              st_pts[:, i, :].T)
        st_pts_n[:, i, :] = _bb(bcc_n, cfs, nCs)
    return st_pts_n


def compose_eval(query, coeff_levels, q_pts, N_c=None, eval_all=False):
    """composite_sphere_diffeomorphism.__call__ / eval_all
    (spherical_spline.py:365-437).  coeff_levels[i] is (3,19,T); applied for
    i = N_c-1 .. 0, each level normalised."""
    if N_c is None:
        N_c = len(coeff_levels)
    q = q_pts.copy()
    outs = [q_pts.copy()]
    for i in reversed(range(N_c)):
        q = map_eval(query, coeff_levels[i], q).T
        outs.append(q)
    return outs if eval_all else q


def scalar_eval(query, coeffs, phi, the):
    """spherical_spline.__call__ (spherical_spline.py:42-70)."""
    ss = np.shape(phi)
    q = np.array(g.sphere2cart(phi.reshape(-1), the.reshape(-1))).T
    return located_eval(query, coeffs, q).reshape(ss)


def lagrange_weights(nVs, t, dt, ts0):
    """Time weights of spline_interp_velocity.__call__ (spherical_spline.py:596-612)."""
    if nVs == 3:
        tau0, tau1, tau2 = ts0, ts0 + dt, ts0 + 2 * dt
        l0 = (t - tau1) * (t - tau2) / (2 * dt ** 2)
        l1 = (t - tau0) * (t - tau2) / (-dt ** 2)
        l2 = (t - tau0) * (t - tau1) / (2 * dt ** 2)
        return (l0, l1, l2)
    if nVs == 2:
        tau = (t - ts0) / dt
        return (1 - tau, tau, 0.0)
    return (1.0, 0.0, 0.0)


class VelocitySpline:
    """spline_interp_velocity (spherical_spline.py:545-655)."""

    def __init__(self, mesh, vals, coeffs, ts):
        self.mesh = mesh
        self.vals = vals          # (3 levels, 3, 4, Nu)
        self.coeffs = coeffs      # (3 levels, 3, 19, Tu)
        self.ts = ts
        self.stepping = False
        self.nVs = 1

    def init_coeffs(self, i):
        self.coeffs[i] = assemble(self.vals[i], self.mesh.simplices, self.mesh.ps_split)

    def __call__(self, t, dt, q_pts):
        outs = located_eval(self.mesh.query, self.coeffs, q_pts.T)   # (3lev, 3, N)
        if self.nVs == 3:
            l0, l1, l2 = lagrange_weights(3, t, dt, self.ts[0])
            return l0 * outs[0] + l1 * outs[1] + l2 * outs[2]
        if self.nVs == 1:
            return outs[0]
        if self.nVs == 2:
            tau = (t - self.ts[0]) / dt
            return (1 - tau) * outs[0] + tau * outs[1]
        return outs


def grad_eval(query, coeffs, q_pts):
    """spherical_spline.gradient (spherical_spline.py:73-111) for a scalar
    spline: tangential gradient of the BB polynomial (d_Proj of the Cartesian
    gradient through bary derivatives)."""
    bcc, trangs, v_pts = query(q_pts)
    bb = bary_minmax(bcc)
    nCs = Cs[bb[0], bb[1]]
    v_pts_n = new_vpoints(v_pts, bb)
    cfs = coeffs[:, trangs]
    n = len(q_pts)
    unos = np.ones([n, 3])
    xs, ys, zs = unos * np.array([1, 0, 0]), unos * np.array([0, 1, 0]), unos * np.array([0, 0, 1])
    V0, V1, V2 = v_pts_n[:, 0, :], v_pts_n[:, 1, :], v_pts_n[:, 2, :]
    bcc_n = g.bary_coords(V0, V1, V2, q_pts)
    bx, by, bz = (g.bary_coords(V0, V1, V2, e) for e in (xs, ys, zs))
    r = np.arange(n)
    c = lambda j: cfs[nCs[:, j], r]
    dp1 = 2 * c(0) * bcc_n[:, 0] + 2 * c(3) * bcc_n[:, 1] + 2 * c(5) * bcc_n[:, 2]
    dp2 = 2 * c(1) * bcc_n[:, 1] + 2 * c(3) * bcc_n[:, 0] + 2 * c(4) * bcc_n[:, 2]
    dp3 = 2 * c(2) * bcc_n[:, 2] + 2 * c(4) * bcc_n[:, 1] + 2 * c(5) * bcc_n[:, 0]
    A = [b[:, 0] * dp1 + b[:, 1] * dp2 + b[:, 2] * dp3 for b in (bx, by, bz)]
    X = q_pts
    outx = (1 - X[:, 0] ** 2) * A[0] + (-X[:, 0] * X[:, 1]) * A[1] + (-X[:, 0] * X[:, 2]) * A[2]
    outy = (-X[:, 0] * X[:, 1]) * A[0] + (1 - X[:, 1] ** 2) * A[1] + (-X[:, 1] * X[:, 2]) * A[2]
    outz = (-X[:, 0] * X[:, 2]) * A[0] + (-X[:, 1] * X[:, 2]) * A[1] + (1 - X[:, 2] ** 2) * A[2]
    return [outx, outy, outz]


def _ambient_grads(query, coeffs, q_pts):
    """Raw values (C,N) and ambient gradients (C,3,N) of the located BB
    polynomials of coeffs (C,19,T): grad p = sum_i dp/db_i grad b_i with the
    barycentric gradients bary_coords(.., e_x / e_y / e_z) of
    spherical_spline.py:228-240 and grad_HBB (700-711)."""
    bcc, trangs, v_pts = query(q_pts)
    bb = bary_minmax(bcc)
    nCs = Cs[bb[0], bb[1]]
    v_pts_n = new_vpoints(v_pts, bb)
    n = len(q_pts)
    unos = np.ones([n, 3])
    V0, V1, V2 = v_pts_n[:, 0, :], v_pts_n[:, 1, :], v_pts_n[:, 2, :]
    bcc_n = g.bary_coords(V0, V1, V2, q_pts)
    Bs = [g.bary_coords(V0, V1, V2, unos * e) for e in np.eye(3)]
    r = np.arange(n)
    vals, grads = [], []
    for cf in coeffs:
        cfs = cf[:, trangs]
        c = lambda j: cfs[nCs[:, j], r]
        vals.append(_bb(bcc_n, cfs[None], nCs)[0])
        dp1 = 2 * c(0) * bcc_n[:, 0] + 2 * c(3) * bcc_n[:, 1] + 2 * c(5) * bcc_n[:, 2]
        dp2 = 2 * c(1) * bcc_n[:, 1] + 2 * c(3) * bcc_n[:, 0] + 2 * c(4) * bcc_n[:, 2]
        dp3 = 2 * c(2) * bcc_n[:, 2] + 2 * c(4) * bcc_n[:, 1] + 2 * c(5) * bcc_n[:, 0]
        grads.append(np.array([b[:, 0] * dp1 + b[:, 1] * dp2 + b[:, 2] * dp3 for b in Bs]))
    return np.array(vals), np.array(grads)


def map_grad(query, coeffs, q_pts):
    """Differential of the normalised map (sphere_diffeomorphism.eval_grad,
    spherical_spline.py:208-265): J[c][d] = grad(X_c / |X|)_d restricted to
    the tangent plane at q, i.e. (I/|X| - X X^T/|X|^3) G (I - q q^T).
    coeffs (3,19,T), q_pts (N,3) -> (3,3,N)."""
    X, G = _ambient_grads(query, coeffs, q_pts)
    r = np.sqrt((X ** 2).sum(0))
    xg = np.einsum("kn,kdn->dn", X, G)
    A = G / r - X[:, None, :] * xg[None, :, :] / r ** 3
    q = q_pts.T
    Aq = np.einsum("cdn,dn->cn", A, q)
    return A - Aq[:, None, :] * q[None, :, :]
