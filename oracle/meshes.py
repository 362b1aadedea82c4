"""Meshes and point location (oracle).  Restates cmm_s2/core/mesh_functions.py.

* ``icosahedral_points`` -- an own icosahedral point generator standing in
  for stripy (the reference's mesh source, absent here): base icosahedron with
  the north pole first and the south pole last (the layout
  mesh_functions.py:137-142 assumes), refined by normalised edge midpoints.
* ``SphericalTriangulation`` -- mesh_functions.py:15-166 with
  refinement_levels=0: simplices from scipy's 3-D Delaunay of {origin} U pts
  (45-56), location by ``find_simplex(q/2)`` (68-88).
* ``StructuredTriangulation`` -- mesh_functions.py:169-318, 352-412.

Pinned by tests/test_oracle_meshes.py: find_simplex location against brute
force containment (all three orientation determinants >= 0) on small meshes;
Delaunay triangle set == midpoint subdivision; ps_split closed forms (unit
tangent vectors, h = sum g_i P_i with the zero third component);
stencil points on the unit sphere at distance ~sqrt(2)*eps; structured-mesh
location checked by brute-force containment (min barycentric >= -1e-12)
and by numpy's own ``//``/``%``/``heaviside`` (same library, bit-identical).
"""
import numpy as np
from scipy.spatial import Delaunay, cKDTree

from . import geometry as g


# ----------------------------------------------------------------------------
# icosahedral point sets (input data; stripy's role in the reference)
# ----------------------------------------------------------------------------
def icosahedron():
    """12 vertices (N pole first, S pole last) and 20 CCW faces."""
    z = 1.0 / np.sqrt(5.0)
    r = 2.0 / np.sqrt(5.0)
    v = [[0.0, 0.0, 1.0]]
    for k in range(5):
        a = 2 * np.pi * k / 5
        v.append([r * np.cos(a), r * np.sin(a), z])
    for k in range(5):
        a = 2 * np.pi * k / 5 + np.pi / 5
        v.append([r * np.cos(a), r * np.sin(a), -z])
    v.append([0.0, 0.0, -1.0])
    v = np.array(v)
    f = []
    for k in range(5):
        u0, u1 = 1 + k, 1 + (k + 1) % 5
        l0, l1 = 6 + k, 6 + (k + 1) % 5
        f.append([0, u0, u1])
        f.append([u0, l0, u1])
        f.append([u1, l0, l1])
        f.append([11, l1, l0])
    f = np.array(f)
    d = g.det_vec([v[f[:, 0]], v[f[:, 1]], v[f[:, 2]]])
    bad = d < 0
    f[bad, 1], f[bad, 2] = f[bad, 2].copy(), f[bad, 1].copy()
    return v, f


def icosahedral_points(k):
    """Vertices of the level-k midpoint-refined icosahedron: N pole at index 0,
    S pole at the last index (mesh_functions.py:137-142 layout)."""
    v, f = icosahedron()
    verts = [p for p in v[:11]]          # keep S pole for the end
    spole = v[11]
    remap = {i: i for i in range(11)}
    remap[11] = -1                       # placeholder index for the S pole
    faces = [[remap[a] for a in t] for t in f]
    pts = list(verts)

    def P(i):
        return spole if i == -1 else pts[i]

    for _ in range(k):
        mids = {}
        nf = []
        for (a, b, c) in faces:
            m = []
            for (i, j) in ((a, b), (b, c), (c, a)):
                key = (min(i, j), max(i, j))
                if key not in mids:
                    x = P(i) + P(j)
                    mids[key] = len(pts)
                    pts.append(x / np.linalg.norm(x))
                m.append(mids[key])
            mab, mbc, mca = m
            nf += [[a, mab, mca], [mab, b, mbc], [mca, mbc, c], [mab, mbc, mca]]
        faces = nf
    pts.append(spole)
    return np.array(pts)


# ----------------------------------------------------------------------------
# icosahedral (unstructured) spherical triangulation
# ----------------------------------------------------------------------------
class SphericalTriangulation:
    """mesh_functions.py:15-166 with the stripy step replaced by given points.

    ``qhull_options`` None reproduces the reference call (41-43); "QJ" is the
    documented substitute for >= ico-6 where default options take hours.
    """

    def __init__(self, pts, qhull_options=None):
        pts = np.asarray(pts, dtype=np.float64)
        self.vertices = pts
        origin = np.array([0, 0, 0])
        ppoints = np.vstack((origin, pts))
        if qhull_options is None:
            mesh_d = Delaunay(ppoints)
        else:
            mesh_d = Delaunay(ppoints, qhull_options=qhull_options)
        simps0 = mesh_d.simplices - 1                        # 45
        simps = np.array([T[T != -1] for T in simps0])       # 47
        v_pts2 = self.vertices[simps]
        dets2 = g.det_vec([v_pts2[:, 0, :], v_pts2[:, 1, :], v_pts2[:, 2, :]])
        inds_1 = np.where(dets2 < 0)                         # 53
        simps[inds_1, 2:0:-1] = simps[inds_1, 1::]           # 56
        self.mesh_d = mesh_d
        self.simplices = simps
        self.x, self.y, self.z = pts[:, 0], pts[:, 1], pts[:, 2]
        self.stencil_pts()
        self.ps_split = ps_split(self.vertices, self.simplices)

    def query(self, q_pts):
        """mesh_functions.py:68-88."""
        trangs = self.mesh_d.find_simplex(q_pts / 2)
        vs = self.vertices[self.simplices[trangs]]
        bcc = g.bary_coords(vs[:, 0, :], vs[:, 1, :], vs[:, 2, :], q_pts)
        return bcc, trangs, vs

    def stencil_pts(self, eps=1e-5):
        """mesh_functions.py:128-166."""
        self.s_pts, self.tan_vects = stencil_points(self.vertices, eps)


def stencil_points(verts, eps=1e-5):
    """mesh_functions.py:128-166 -> s_pts (3,4,Nv), tan_vects [a1_n, a2_n]."""
    a1 = np.array(g.cross(verts.T, [0, 0, 1])).T
    a2 = np.array(g.cross(verts.T, a1.T)).T
    a1[0, :] = g.cross(verts[0, :], [1 / np.sqrt(2), 0, 1 / np.sqrt(2)])
    a1[-1, :] = g.cross(verts[-1, :], [1 / np.sqrt(2), 0, -1 / np.sqrt(2)])
    a2[0, :] = g.cross(verts[0, :], a1[0, :])
    a2[-1, :] = g.cross(verts[-1, :], a1[-1, :])
    a1_n, a2_n = g.div_norm(a1).T, g.div_norm(a2).T
    eps_1p = verts - a1_n * eps - a2_n * eps
    eps_1m = verts + a1_n * eps - a2_n * eps
    eps_2p = verts - a1_n * eps + a2_n * eps
    eps_2m = verts + a1_n * eps + a2_n * eps
    spts = [g.pi_proj(eps_1p, verts), g.pi_proj(eps_1m, verts),
            g.pi_proj(eps_2p, verts), g.pi_proj(eps_2m, verts)]
    spts = [np.array([spts[0][c, :], spts[1][c, :], spts[2][c, :], spts[3][c, :]])
            for c in range(3)]
    return np.array(spts), [a1_n, a2_n]


def ps_split(vertices, simplices):
    """mesh_functions.py:92-126 (identical copy at 284-318)."""
    inds = np.array(simplices)
    v_pts = vertices[inds]
    v1, v2, v3 = v_pts[:, 0, :].copy(), v_pts[:, 1, :].copy(), v_pts[:, 2, :].copy()
    v4 = g.div_norm((v1 + v2 + v3) / 3).T
    e1 = g.div_norm(v1 / 2 + v2 / 2).T
    e2 = g.div_norm(v2 / 2 + v3 / 2).T
    e3 = g.div_norm(v3 / 2 + v1 / 2).T
    stp = g.sphere_tan_proj
    h12, h21, h23 = stp(v2 - v1, v1).T, stp(v1 - v2, v2).T, stp(v3 - v2, v2).T
    h32, h13, h31 = stp(v2 - v3, v3).T, stp(v3 - v1, v1).T, stp(v1 - v3, v3).T
    h41, h42, h43 = stp(v4 - v1, v1).T, stp(v4 - v2, v2).T, stp(v4 - v3, v3).T
    bc = g.bary_coords
    g12, g21 = bc(v1, e1, v4, h12), bc(v2, v4, e1, h21)
    g23, g32 = bc(v2, e2, v4, h23), bc(v3, v4, e2, h32)
    g13, g31 = bc(v1, v4, e3, h13), bc(v3, e3, v4, h31)
    g14 = bc(v1, v4, e3, stp(v4, v1).T)
    g24 = bc(v2, v4, e1, stp(v4, v2).T)
    g34 = bc(v3, v4, e2, stp(v4, v3).T)
    mid = bc(v1, v2, v3, v4)
    e_1, e_2, e_3 = bc(v1, v2, v3, e1), bc(v1, v2, v3, e2), bc(v1, v2, v3, e3)
    Hs = [h12.T, h21.T, h23.T, h32.T, h13.T, h31.T, h41.T, h42.T, h43.T]
    Gs = [g12, g21, g23, g32, g13, g31, g14, g24, g34, mid]
    Es = [e_1, e_2, e_3]
    return Hs, Gs, Es


def locate_kdtree(vertices, simplices, q, k=12):
    """Fast containment search for large tests: k nearest triangle centroids,
    first candidate with all barycentrics >= 0, else the most-inside one.
    Checked against ``SphericalTriangulation.query`` in the oracle tests."""
    cen = vertices[simplices].sum(axis=1)
    cen /= np.linalg.norm(cen, axis=1)[:, None]
    tree = cKDTree(cen)
    _, cand = tree.query(q, k=k)
    best = np.full(len(q), -1, dtype=np.int64)
    bestm = np.full(len(q), -np.inf)
    for c in range(k):
        t = cand[:, c]
        vs = vertices[simplices[t]]
        b = g.bary_coords(vs[:, 0], vs[:, 1], vs[:, 2], q)
        m = b.min(axis=1)
        upd = (bestm < 0) & (m > bestm)
        best[upd] = t[upd]
        bestm[upd] = m[upd]
    return best, bestm


def orientation_margin(vertices, simplices, tri, q):
    """min over the 3 edges of det(v_i, v_j, q) / (|v_i x v_j|) for the given
    triangle -- the normalised orientation margin used by the tie contract."""
    vs = vertices[simplices[tri]]
    out = []
    for (i, j) in ((1, 2), (2, 0), (0, 1)):
        n = np.cross(vs[:, i], vs[:, j])
        out.append(np.einsum("ij,ij->i", n, q) / np.linalg.norm(n, axis=1))
    return np.min(np.stack(out, axis=1), axis=1)


# ----------------------------------------------------------------------------
# MWSS structured velocity mesh
# ----------------------------------------------------------------------------
def mwss_thetas(L):
    """ssht MWSS colatitudes theta_t = 2 t pi / (2 L), t = 0..L."""
    return np.array([2.0 * t * np.pi / (2.0 * L) for t in range(L + 1)])


def mwss_phis(L):
    """ssht MWSS longitudes phi_p = 2 p pi / (2 L), p = 0..2L-1."""
    return np.array([2.0 * p * np.pi / (2.0 * L) for p in range(2 * L)])


def grid_assemble(N, M):
    """mesh_functions.py:380-412 (loop form, kept literal)."""
    out = []
    for i in range(N - 1):
        out.append([0, i + 1, i + 2])
    out.append([0, N, 1])
    for i in range(0, M - 3):
        for j in range(1, N + 1):
            k = j + i * N
            if j != N:
                out.append([k, k + N + 1, k + 1])
                out.append([k, k + N, k + N + 1])
            else:
                out.append([k, k + 1, i * N + 1])
                out.append([k, k + N, k + 1])
    nn = (M - 3) * N + 1
    for i in range(nn, nn + N - 1):
        out.append([i, nn + N, i + 1])
    out.append([i + 1, nn + N, i + 1 - (N - 1)])
    return out


def full_assembly(N, M):
    """mesh_functions.py:352-378."""
    simps = grid_assemble(N, M)
    simplices_p1 = [T + [i] for (T, i) in zip(simps, range(len(simps)))]
    msimplices = np.empty([M - 1, N, 2, 4], dtype=int)
    for j in range(N):
        msimplices[0, j, 0, :] = simplices_p1[j]
        msimplices[0, j, 1, :] = simplices_p1[j]
    k = N
    for i in range(1, M - 2):
        for j in range(N):
            msimplices[i, j, 0, :] = simplices_p1[k]
            msimplices[i, j, 1, :] = simplices_p1[k + 1]
            k += 2
    for j in range(N):
        msimplices[M - 2, j, 0, :] = simplices_p1[j - N]
        msimplices[M - 2, j, 1, :] = simplices_p1[j - N]
    return simps, msimplices


class StructuredTriangulation:
    """mesh_functions.py:169-318.  Vertices: [N pole, rings theta_1..theta_{L-1}
    x phi_0..phi_{2L-1} (row major), S pole]; coordinates from sphere2cart (the
    reference obtains the same points through stripy's lon/lat->xyz; the poles
    are the phi=0 grid samples of rings 0 and L)."""

    def __init__(self, L):
        thetas, phis = mwss_thetas(L), mwss_phis(L)
        Phi, The = np.meshgrid(phis, thetas)
        self.grid_points = g.sphere2cart(Phi, The)
        N, M = len(phis), len(thetas)
        P2, T2 = np.meshgrid(phis, thetas[1:-1])
        ring = np.array(g.sphere2cart(P2.reshape(-1), T2.reshape(-1))).T
        npole = np.array(g.sphere2cart(np.array([0.0]), np.array([thetas[0]]))).T
        spole = np.array(g.sphere2cart(np.array([0.0]), np.array([thetas[-1]]))).T
        self.vertices = np.vstack([npole, ring, spole])
        simplices, msimplices = full_assembly(N, M)
        self.phi = phis
        self.theta = thetas
        self.L = L
        self.simplices = np.array(simplices)
        self.msimplices = msimplices
        self.ps_split = ps_split(self.vertices, self.simplices)

    def query(self, q_pts):
        """mesh_functions.py:196-199."""
        phi, theta = g.cart2sphere(q_pts.T)
        ijs = self.inds(phi, theta)
        return self.containing_simplex_and_bcc_structured(ijs, q_pts)

    def inds(self, phiq, thetaq):
        """mesh_functions.py:201-255."""
        dphi = abs(self.phi[1] - self.phi[0])
        dthe = abs(self.theta[1] - self.theta[0])
        phiq_n = phiq % (2 * np.pi)
        phi_l = ((phiq_n - self.phi[0]) // dphi).astype(int) % (len(self.phi))
        theta_l = ((thetaq - self.theta[0]) // dthe).astype(int)
        theta_l[theta_l == (len(self.theta) - 1)] = len(self.theta) - 2
        theta_c = (thetaq - self.theta[theta_l]) / dthe
        inds0 = np.where((theta_c <= 0.03) & (theta_l != 0))
        inds1 = np.where((theta_c > 0.85) & (theta_l != len(self.theta) - 2))
        phi_p1 = (phi_l + 1) % (len(self.phi))
        the_p1 = theta_l + 1
        v_0 = np.array(g.sphere2cart(self.phi[phi_l[inds0]], self.theta[theta_l[inds0]]))
        v_1 = np.array(g.sphere2cart(self.phi[phi_p1[inds0]], self.theta[theta_l[inds0]]))
        n_vs = g.cross(v_0, v_1)
        q_pts0 = np.array(g.sphere2cart(phiq_n[inds0], thetaq[inds0]))
        s_inds0 = np.heaviside(g.dot(n_vs, q_pts0), 0).astype(int)
        theta_l[inds0] = theta_l[inds0] - s_inds0
        v_01 = np.array(g.sphere2cart(self.phi[phi_l[inds1]], self.theta[the_p1[inds1]]))
        v_11 = np.array(g.sphere2cart(self.phi[phi_p1[inds1]], self.theta[the_p1[inds1]]))
        n_vs2 = g.cross(v_01, v_11)
        q_pts1 = np.array(g.sphere2cart(phiq_n[inds1], thetaq[inds1]))
        s_inds1 = np.heaviside(-g.dot(n_vs2, q_pts1), 0).astype(int)
        theta_l[inds1] = theta_l[inds1] + s_inds1
        return [phi_l, theta_l]

    def containing_simplex_and_bcc_structured(self, inds, pts):
        """mesh_functions.py:257-282."""
        phi0, theta0 = self.phi[inds[0]], self.theta[inds[1]]
        v_0 = g.sphere2cart(phi0, theta0)
        ijsp1 = [inds[0].copy() + 1, inds[1].copy() + 1]
        ijsp1[0] = ijsp1[0] % (len(self.phi))
        phi_p1, theta_p1 = self.phi[ijsp1[0]], self.theta[ijsp1[1]]
        v_1 = np.array(g.sphere2cart(phi_p1, theta_p1))
        n_vs = g.cross(v_0, v_1)
        s_inds = np.heaviside(-g.dot(n_vs, pts.T), 0).astype(int)
        tri_out_temp = self.msimplices[inds[1], inds[0], :]
        l = np.shape(tri_out_temp)[0]
        tri_out = tri_out_temp[np.arange(0, l), s_inds]
        trangs = tri_out[:, 3]
        verts = self.vertices[tri_out[:, 0:3]]
        bcc = g.bary_coords(verts[:, 0, :], verts[:, 1, :], verts[:, 2, :], pts)
        return bcc, trangs, verts

    def sample_points(self):
        """(N,3) MWSS sample grid in row-major [theta, phi] order
        (spherical_simulations.py:210-214)."""
        return np.array([self.grid_points[c].reshape(-1) for c in range(3)]).T
