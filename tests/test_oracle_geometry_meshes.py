"""CPU: pins of the oracle's geometry and mesh restatements
(cmm_s2/core/utils.py, mesh_functions.py) by closed forms, brute force and
independent library routines."""
import numpy as np
import pytest

from helpers import random_unit
from oracle import geometry as g, meshes as om


def test_bary_coords_reconstruct_point():
    rng = np.random.default_rng(0)
    v = [random_unit(500, s) for s in (1, 2, 3)]
    q = random_unit(500, 4)
    b = g.bary_coords(v[0], v[1], v[2], q)
    rec = b[:, :1] * v[0] + b[:, 1:2] * v[1] + b[:, 2:] * v[2]
    assert np.abs(rec - q).max() < 1e-10
    # det_vec against numpy's LU determinant (columns a, b, c)
    d = g.det_vec(v)
    ref = np.linalg.det(np.stack(v, axis=2))
    assert np.abs(d - ref).max() < 1e-13


def test_sphere_maps_and_projections():
    q = random_unit(1000, 5)
    phi, th = g.cart2sphere(q.T)
    assert ((phi >= 0) & (phi < 2 * np.pi)).all() and ((th >= 0) & (th <= np.pi)).all()
    back = np.array(g.sphere2cart(phi, th)).T
    assert np.abs(back - q).max() < 4e-15
    # pi_proj lifts tangent-plane points back to the unit sphere along n
    y = q + 1e-3 * np.cross(q, random_unit(1000, 6))
    p = g.pi_proj(y, q).T
    assert np.abs(np.linalg.norm(p, axis=1) - 1).max() < 1e-15
    # sphere_tan_proj: unit length and tangent at the base point
    a = random_unit(1000, 7) - q
    h = g.sphere_tan_proj(a, q).T
    # the reference's d = sqrt(a.a - 2c^2 + (b.b)c^2) cancels when a is close
    # to the base point's direction
    assert np.abs(np.linalg.norm(h, axis=1) - 1).max() < 1e-12
    assert np.abs(np.einsum("ij,ij->i", h, q)).max() < 1e-12
    # tan_proj is I - x x^T
    A = random_unit(1000, 8).T
    t = np.array(g.tan_proj(q.T, A))
    ref = A - np.einsum("ij,ij->j", q.T, A) * q.T
    assert np.abs(t - ref).max() < 1e-15


@pytest.mark.parametrize("k", [0, 1, 2, 3])
def test_icosahedral_points_and_delaunay(k):
    pts = om.icosahedral_points(k)
    assert len(pts) == 10 * 4 ** k + 2
    assert np.abs(np.linalg.norm(pts, axis=1) - 1).max() < 1e-15
    assert np.allclose(pts[0], [0, 0, 1]) and np.allclose(pts[-1], [0, 0, -1])
    m = om.SphericalTriangulation(pts)
    assert len(m.simplices) == 20 * 4 ** k
    v = pts[m.simplices]
    assert (g.det_vec([v[:, 0], v[:, 1], v[:, 2]]) > 0).all()
    # the Delaunay triangle set equals the midpoint subdivision of the product
    from cmm_s2.core import _meshgen as G
    ico = G.IcosahedralMesh(k)
    key = lambda S, P: {tuple(sorted(map(tuple, np.round(P[t], 12)))) for t in S}
    assert key(m.simplices, pts) == key(ico.simplices, ico.points)


def test_find_simplex_location_against_brute_force():
    pts = om.icosahedral_points(2)
    m = om.SphericalTriangulation(pts)
    q = np.concatenate([random_unit(3000, 9), pts])
    bcc, tri, vs = m.query(q)
    assert (tri >= 0).all()
    v = pts[m.simplices]
    # brute force: every triangle with all three orientation determinants >= 0
    for i in range(0, len(q), 97):
        d = np.stack([g.det_vec([np.repeat(q[i:i + 1], len(v), 0), v[:, 1], v[:, 2]]),
                      g.det_vec([v[:, 0], np.repeat(q[i:i + 1], len(v), 0), v[:, 2]]),
                      g.det_vec([v[:, 0], v[:, 1], np.repeat(q[i:i + 1], len(v), 0)])], 1)
        inside = np.where(d.min(axis=1) >= -1e-14)[0]
        assert tri[i] in inside
    rec = np.einsum("ni,nij->nj", bcc, vs)
    assert np.abs(rec - q).max() < 1e-14
    # the fast containment search agrees wherever the triangle is unique
    t2, mm = om.locate_kdtree(pts, m.simplices, q)
    clear = om.orientation_margin(pts, m.simplices, tri, q) > 1e-12
    assert (t2[clear] == tri[clear]).all() and (mm > -1e-14).all()


def test_ps_split_closed_forms():
    pts = om.icosahedral_points(2)
    m = om.SphericalTriangulation(pts)
    Hs, Gs, Es = m.ps_split
    v = pts[m.simplices]
    v1, v2, v3 = v[:, 0], v[:, 1], v[:, 2]
    base = [v1, v2, v2, v3, v1, v3, v1, v2, v3]
    for H, b in zip(Hs, base):
        H = H.T
        assert np.abs(np.linalg.norm(H, axis=1) - 1).max() < 1e-14      # unit
        assert np.abs(np.einsum("ij,ij->i", H, b)).max() < 1e-14        # tangent at its vertex
    nrm = lambda x: x / np.linalg.norm(x, axis=1)[:, None]
    e1, e2, e3, v4 = nrm(v1 + v2), nrm(v2 + v3), nrm(v3 + v1), nrm(v1 + v2 + v3)
    tris = [(v1, e1, v4), (v2, v4, e1), (v2, e2, v4), (v3, v4, e2), (v1, v4, e3), (v3, e3, v4)]
    zero_col = [2, 1, 2, 1, 1, 2]   # the split vertex v4 carries no weight: h lies in plane(v_i, e)
    for gi, (P1, P2, P3) in enumerate(tris):
        G = Gs[gi]
        rec = G[:, :1] * P1 + G[:, 1:2] * P2 + G[:, 2:] * P3
        assert np.abs(rec - Hs[gi].T).max() < 1e-13
        assert np.abs(G[:, zero_col[gi]]).max() < 1e-12
    # edge midpoints have macro barycentrics (1/|vi+vj|) on their two vertices
    E1 = Es[0]
    assert np.abs(E1[:, 2]).max() < 1e-13
    assert np.abs(E1[:, 0] - 1 / np.linalg.norm(v1 + v2, axis=1)).max() < 1e-13


def test_stencil_points():
    pts = om.icosahedral_points(3)
    s, (a1, a2) = om.stencil_points(pts)
    assert s.shape == (3, 4, len(pts))
    S = np.transpose(s, (2, 1, 0))
    assert np.abs(np.linalg.norm(S, axis=2) - 1).max() < 1e-15
    d = np.linalg.norm(S - pts[:, None, :], axis=2)
    assert np.abs(d - np.sqrt(2) * 1e-5).max() < 1e-12
    # orthonormal tangent frames, stencil k at (-+a1, -+a2) in the order of
    # mesh_functions.py:148-160
    for a in (a1, a2):
        assert np.abs(np.linalg.norm(a, axis=1) - 1).max() < 1e-15
        assert np.abs(np.einsum("ij,ij->i", a, pts)).max() < 1e-15
    assert np.abs(np.einsum("ij,ij->i", a1, a2)).max() < 1e-15
    sgn = [(-1, -1), (1, -1), (-1, 1), (1, 1)]
    for k, (s1, s2) in enumerate(sgn):
        dv = S[:, k] - pts
        assert (np.sign(np.einsum("ij,ij->i", dv, a1)) == s1).all()
        assert (np.sign(np.einsum("ij,ij->i", dv, a2)) == s2).all()


@pytest.mark.parametrize("L", [4, 8, 16])
def test_structured_mesh(L):
    from cmm_s2.core import _meshgen as G
    o = om.StructuredTriangulation(L)
    assert len(o.vertices) == 2 * L * (L - 1) + 2 and len(o.simplices) == 4 * L * (L - 1)
    # closed forms of the product equal the reference's loops restated
    assert np.array_equal(G.mwss_simplices(L), o.simplices)
    assert np.array_equal(G.mwss_msimplices(L), o.msimplices)
    v = o.vertices[o.simplices]
    assert (g.det_vec([v[:, 0], v[:, 1], v[:, 2]]) > 0).all()
    # location: the returned triangle contains the point (brute-force check
    # with homogeneous barycentrics) for random points and grid points
    q = np.concatenate([random_unit(20000, L), o.sample_points()])
    bcc, tri, vs = o.query(q)
    miss = bcc.min(axis=1) < -1e-12
    if L >= 8:
        assert not miss.any()
    else:
        # reference quirk: at very coarse bandlimits the 0.03 / 0.85 band
        # corrections of mesh_functions.py:229-253 do not cover the
        # great-circle bulge of the ring edges, and a few points get a
        # neighbouring triangle (restated as is; the GPU does the same)
        assert 0 < miss.mean() < 0.01
    rec = np.einsum("ni,nij->nj", bcc, vs)
    assert np.abs(rec - q).max() < 1e-13


def test_numpy_floor_division_semantics_restated_on_gpu_side():
    """The GPU locator restates numpy's floor division (npy_divmod), which is
    not floor(a / b): 1.0 // 0.1 == 9."""
    from cmm_s2.core import _meshgen  # noqa: F401  (import check only)
    assert np.floor_divide(1.0, 0.1) == 9.0 and np.floor(1.0 / 0.1) == 10.0
