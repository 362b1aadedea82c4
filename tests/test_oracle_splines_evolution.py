"""CPU: pins of the oracle's spline, time-stepping and composition
restatements (spherical_spline.py, evolution_functions.py)."""
import numpy as np
import pytest

from helpers import perturbed_identity_jets, random_unit
from oracle import evolution as oev, geometry as g, meshes as om, splines as osp


def _quadratic_jets(V, A, comps=3):
    """Jets of p_c(x) = x^T A_c x: vals[c,0] = p_c, vals[d,1+c] = d_d p_c
    (ambient gradient; only h . grad with tangent h enters the formulas)."""
    jets = np.zeros((comps, 4, len(V)))
    for c in range(comps):
        jets[c, 0] = np.einsum("ni,ij,nj->n", V, A[c], V)
        jets[:, 1 + c] = (2 * V @ A[c]).T
    return jets


@pytest.mark.parametrize("mesh_kind", ["ico", "mwss"])
def test_exact_reproduction_of_homogeneous_quadratics(mesh_kind):
    """The PS formulas reproduce any homogeneous quadratic exactly: c3 = P(v1,e1),
    c12 = p(e1), c15 = P(e1,v4), c18 = p(v4) (polar forms).  A wrong index in
    Cs / edges_ps or a wrong coefficient formula breaks this."""
    rng = np.random.default_rng(1)
    A = rng.standard_normal((3, 3, 3))
    A = A + A.transpose(0, 2, 1)
    if mesh_kind == "ico":
        m = om.SphericalTriangulation(om.icosahedral_points(2))
    else:
        m = om.StructuredTriangulation(8)
    c = osp.assemble(_quadratic_jets(m.vertices, A), m.simplices, m.ps_split)
    q = random_unit(20000, 2)
    raw = osp.located_eval(m.query, c, q)
    exact = np.array([np.einsum("ni,ij,nj->n", q, A[k], q) for k in range(3)])
    assert np.abs(raw - exact).max() < 1e-12
    # stencil_eval extrapolates the located sub-triangle's polynomial: still exact
    if mesh_kind == "ico":
        yv = q[:m.vertices.shape[0]].T.copy()
        st = np.stack([yv + 0.01 * random_unit(len(m.vertices), 10 + k).T for k in range(4)], 1)
        sv = osp.stencil_eval(m.query, c, yv, st.copy())
        for k in range(4):
            ex = np.array([np.einsum("in,ij,jn->n", st[:, k], A[cc], st[:, k]) for cc in range(3)])
            assert np.abs(sv[:, k] - ex).max() < 1e-12


def test_vertex_interpolation_and_continuity():
    m = om.SphericalTriangulation(om.icosahedral_points(2))
    jets = perturbed_identity_jets(m.vertices, 3)
    c = osp.assemble(jets, m.simplices, m.ps_split)
    raw = osp.located_eval(m.query, c, m.vertices)
    assert np.abs(raw - jets[:, 0]).max() < 1e-13
    # C0 across macro edges: evaluate the two triangles sharing an edge at the
    # edge midpoint with their own coefficients
    S = m.simplices
    e2t = {}
    for t, tri in enumerate(S):
        for i in range(3):
            e2t.setdefault(tuple(sorted((tri[i], tri[(i + 1) % 3]))), []).append(t)
    V = m.vertices
    worst = 0.0
    for (a, b), ts in list(e2t.items())[:200]:
        p = V[a] + V[b] + 0.3 * (V[a] - V[b])
        p /= np.linalg.norm(p)
        vals = []
        for t in ts:
            vs = V[S[t]][None]
            bc = g.bary_coords(vs[:, 0], vs[:, 1], vs[:, 2], p[None])
            bb = osp.bary_minmax(bc)
            vn = osp.new_vpoints(vs, bb)
            bn = g.bary_coords(vn[:, 0], vn[:, 1], vn[:, 2], p[None])
            vals.append(osp._bb(bn, c[:, :, [t]], osp.Cs[bb[0], bb[1]]))
        worst = max(worst, np.abs(vals[0] - vals[1]).max())
    assert worst < 1e-13


def test_convergence_order_on_reference_test_function():
    """f = x^5 + xyz + sin y (reference tests/test_spline_interpolant.py:40-47):
    scalar PS interpolation on MWSS meshes converges at order 3."""
    def f(X):
        return X[:, 0] ** 5 + X[:, 0] * X[:, 1] * X[:, 2] + np.sin(X[:, 1])

    def grad(X):
        gx = 5 * X[:, 0] ** 4 + X[:, 2] * X[:, 1]
        gy = X[:, 2] * X[:, 0] + np.cos(X[:, 1])
        gz = X[:, 1] * X[:, 0]
        G = np.stack([gx, gy, gz], 1)
        return G - np.einsum("ij,ij->i", G, X)[:, None] * X

    q = random_unit(20000, 3)
    errs = []
    for L in (8, 16, 32):
        m = om.StructuredTriangulation(L)
        vals = np.zeros((4, len(m.vertices)))
        vals[0] = f(m.vertices)
        vals[1:] = grad(m.vertices).T
        c = osp.assemble(vals, m.simplices, m.ps_split)
        errs.append(np.abs(osp.located_eval(m.query, c, q) - f(q)).max())
    orders = np.log2(np.array(errs[:-1]) / np.array(errs[1:]))
    assert (orders > 2.7).all(), orders


def test_lagrange_weights_reproduce_quadratics_in_time():
    for t in (0.37, 0.2, 0.31):
        w = osp.lagrange_weights(3, t, 0.1, 0.1)
        taus = np.array([0.1, 0.2, 0.3])
        for p in (lambda s: 1 + 0 * s, lambda s: s, lambda s: 3 * s * s - s):
            assert abs(np.dot(w, p(taus)) - p(t)) < 1e-13
    assert osp.lagrange_weights(2, 0.15, 0.1, 0.1) == pytest.approx((0.5, 0.5, 0.0))


def _rot_velocity(axis, rate=2 * np.pi):
    ax = np.asarray(axis, float) / np.linalg.norm(axis)
    return lambda t, dt, X: rate * np.cross(ax[:, None], X, axis=0)


@pytest.mark.parametrize("name,order", [("Euler_step_proj", 1), ("improved_Euler_proj", 2),
                                        ("RK4_proj", 4), ("RK2_proj", 2)])
def test_integrator_orders_on_solid_body_rotation(name, order):
    """Backward steps of u = w a x X: exact foot is the rotation by -w dt."""
    V = _rot_velocity([0.2, -0.4, 1.0])
    x = random_unit(200, 4).T
    errs = []
    for dt in (0.02, 0.01):
        y = getattr(oev, name)(0.0, dt, x, V)
        R = g.rotation_matrix([0.2, -0.4, 1.0], -2 * np.pi * dt)
        errs.append(np.abs(y - R @ x).max())
    ratio = errs[0] / errs[1]
    assert abs(np.log2(ratio) - (order + 1)) < 0.3, (name, ratio)


def test_spline_projection_of_a_rotation():
    """spline_proj_sphere of exact stencil feet of a rotation R: value R v and
    transposed Jacobian (R (I - v v^T))^T up to the eps^2 FD error."""
    m = om.SphericalTriangulation(om.icosahedral_points(3))
    R = g.rotation_matrix([1.0, 0.5, -0.3], 0.2)
    st = np.einsum("ij,jkn->ikn", R, m.s_pts)

    class It:
        pass
    it = It()
    it.vals = np.zeros((3, 4, len(m.vertices)))
    it.mesh = m
    mp = oev.spline_proj_sphere(it, st)
    V = m.vertices
    assert np.abs(mp.vals[:, 0] - R @ V.T).max() < 1e-10
    P = np.eye(3)[None] - np.einsum("ni,nj->nij", V, V)
    J = np.einsum("ij,njk->nik", R, P)      # J[n, c, d] = d X_c / d x_d
    got = mp.vals[:, 1:]                    # got[d, c, n]
    assert np.abs(got - np.transpose(J, (2, 1, 0))).max() < 1e-9


def test_compose_of_rotation_maps():
    m = om.SphericalTriangulation(om.icosahedral_points(3))
    Rs = [g.rotation_matrix([1, 2, 3], 0.1), g.rotation_matrix([-1, 0, 2], 0.2)]
    V = m.vertices
    P = np.eye(3)[None] - np.einsum("ni,nj->nij", V, V)
    levels = []
    for R in Rs:
        jets = np.zeros((3, 4, len(V)))
        jets[:, 0] = R @ V.T
        J = np.einsum("ij,njk->nik", R, P)
        jets[:, 1:] = np.transpose(J, (2, 1, 0))
        levels.append(osp.assemble(jets, m.simplices, m.ps_split))
    q = random_unit(3000, 5)
    out = osp.compose_eval(m.query, levels, q)      # level 1 first, then level 0
    exact = (Rs[0] @ Rs[1] @ q.T).T
    assert np.abs(out - exact).max() < 2e-3     # spline error O(h^3) of a rotation on ico-3
    out4 = osp.compose_eval(om.SphericalTriangulation(om.icosahedral_points(4)).query,
                            [osp.assemble(_rot_jets(R, om.icosahedral_points(4)),
                                          om.SphericalTriangulation(om.icosahedral_points(4)).simplices,
                                          om.SphericalTriangulation(om.icosahedral_points(4)).ps_split)
                             for R in Rs], q)
    assert np.abs(out4 - exact).max() < np.abs(out - exact).max() / 4


def _rot_jets(R, V):
    P = np.eye(3)[None] - np.einsum("ni,nj->nij", V, V)
    jets = np.zeros((3, 4, len(V)))
    jets[:, 0] = R @ V.T
    jets[:, 1:] = np.transpose(np.einsum("ij,njk->nik", R, P), (2, 1, 0))
    return jets


def _tangent_fd(f, q, h=1e-6):
    """Central differences of f along the two tangent directions at q (N,3):
    returns (J t1, J t2) estimates and the directions."""
    a = np.cross(q, np.array([0.3, -0.5, 0.8]))
    t1 = a / np.linalg.norm(a, axis=1)[:, None]
    t2 = np.cross(q, t1)
    out = []
    for t in (t1, t2):
        qp = q + h * t
        qm = q - h * t
        qp /= np.linalg.norm(qp, axis=1)[:, None]
        qm /= np.linalg.norm(qm, axis=1)[:, None]
        out.append((f(qp) - f(qm)) / (2 * h))
    return out, (t1, t2)


def test_map_jacobian_against_finite_differences_and_quadratics():
    m = om.SphericalTriangulation(om.icosahedral_points(3))
    jets = perturbed_identity_jets(m.vertices, 7, amp=0.05)
    c = osp.assemble(jets, m.simplices, m.ps_split)
    q = random_unit(3000, 11)
    J = osp.map_grad(m.query, c, q)
    (d1, d2), (t1, t2) = _tangent_fd(lambda x: osp.map_eval(m.query, c, x), q)
    for d, t in ((d1, t1), (d2, t2)):
        Jt = np.einsum("cdn,nd->cn", J, t)
        assert np.abs(Jt - d).max() < 1e-5 * np.abs(d).max()
    # exact for a homogeneous quadratic map p(x) = (x^T A_c x)_c: the spline
    # reproduces p, so the Jacobian is the analytic one
    rng = np.random.default_rng(2)
    A = rng.standard_normal((3, 3, 3))
    A = A + A.transpose(0, 2, 1) + 6 * np.eye(3)[None]
    c2 = osp.assemble(_quadratic_jets(m.vertices, A), m.simplices, m.ps_split)
    X = np.array([np.einsum("ni,ij,nj->n", q, A[k], q) for k in range(3)])
    G = np.array([(2 * q @ A[k]).T for k in range(3)])
    r = np.sqrt((X ** 2).sum(0))
    xg = np.einsum("kn,kdn->dn", X, G)
    Aj = G / r - X[:, None] * xg[None] / r ** 3
    ref = Aj - np.einsum("cdn,dn->cn", Aj, q.T)[:, None] * q.T[None]
    assert np.abs(osp.map_grad(m.query, c2, q) - ref).max() < 1e-11 * np.abs(ref).max()
    # tangent at both ends: J q = 0 and X^T J = 0
    assert np.abs(np.einsum("cdn,nd->cn", J, q)).max() < 1e-12
    Xn = osp.map_eval(m.query, c, q)
    assert np.abs(np.einsum("cn,cdn->dn", Xn, J)).max() < 1e-12


def test_scalar_gradient_against_finite_differences():
    m = om.StructuredTriangulation(16)
    V = m.vertices
    f = lambda X: X[:, 0] ** 3 + X[:, 1] * X[:, 2] + np.sin(2 * X[:, 2])
    gr = np.stack([3 * V[:, 0] ** 2, V[:, 2], V[:, 1] + 2 * np.cos(2 * V[:, 2])], 1)
    gr = gr - np.einsum("ij,ij->i", gr, V)[:, None] * V
    vals = np.zeros((4, len(V)))
    vals[0] = f(V)
    vals[1:] = gr.T
    c = osp.assemble(vals, m.simplices, m.ps_split)
    q = random_unit(2000, 12)
    grad = np.array(osp.grad_eval(m.query, c, q))
    (d1, d2), (t1, t2) = _tangent_fd(lambda x: osp.located_eval(m.query, c, x), q)
    for d, t in ((d1, t1), (d2, t2)):
        assert np.abs(np.einsum("dn,nd->n", grad, t) - d).max() < 1e-5 * np.abs(d).max()
