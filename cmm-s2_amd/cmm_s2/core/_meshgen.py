"""Host-side mesh generation (setup, not the hot path): icosahedral point sets,
convex-hull triangulation of arbitrary sphere points, midpoint refinement and
the closed-form MWSS structured triangulation."""
import numpy as np


def _det(a, b, c):
    return np.einsum("ij,ij->i", a, np.cross(b, c))


def orient(vertices, simplices):
    """Swap the last two vertices of clockwise triangles
    (mesh_functions.py:49-56)."""
    s = np.array(simplices, dtype=np.int64, copy=True)
    v = vertices[s]
    bad = _det(v[:, 0], v[:, 1], v[:, 2]) < 0
    s[bad, 1], s[bad, 2] = s[bad, 2].copy(), s[bad, 1].copy()
    return s


def hull_triangles(pts):
    """Triangles of the convex hull of points on the unit sphere (the
    spherical Delaunay triangulation the reference takes from stripy / Qhull)."""
    from scipy.spatial import ConvexHull
    return orient(pts, ConvexHull(pts).simplices)


def delaunay_triangles(pts, qhull_options=None):
    """The reference's triangulation and row order (mesh_functions.py:41-56):
    Qhull's 3-D Delaunay tessellation of the origin and the points, the
    origin dropped from every tetrahedron (the remaining three indices kept
    in their order), clockwise rows fixed by swapping their last two
    vertices.  Returns (simplices, the Delaunay object)."""
    from scipy.spatial import Delaunay
    pp = np.vstack((np.zeros((1, 3)), pts))
    d = Delaunay(pp) if qhull_options is None else Delaunay(pp, qhull_options=qhull_options)
    s0 = d.simplices - 1
    s = np.array([t[t != -1] for t in s0], dtype=np.int64)
    return orient(pts, s), d


def refine(vertices, simplices, levels):
    """Midpoint bisection: every triangle (a,b,c) -> (a,ab,ca), (ab,b,bc),
    (ca,bc,c), (ab,bc,ca) with normalised edge midpoints appended after the
    existing vertices."""
    V = np.asarray(vertices, dtype=np.float64)
    S = np.asarray(simplices, dtype=np.int64)
    for _ in range(levels):
        e = np.concatenate([S[:, [0, 1]], S[:, [1, 2]], S[:, [2, 0]]])
        e.sort(axis=1)
        uniq, inv = np.unique(e, axis=0, return_inverse=True)
        inv = inv.reshape(-1)
        mid = V[uniq[:, 0]] + V[uniq[:, 1]]
        mid /= np.linalg.norm(mid, axis=1)[:, None]
        base = len(V)
        V = np.vstack([V, mid])
        T = len(S)
        mab, mbc, mca = inv[:T] + base, inv[T:2 * T] + base, inv[2 * T:] + base
        a, b, c = S[:, 0], S[:, 1], S[:, 2]
        S = np.stack([np.stack([a, mab, mca], 1), np.stack([mab, b, mbc], 1),
                      np.stack([mca, mbc, c], 1), np.stack([mab, mbc, mca], 1)], 1).reshape(-1, 3)
    return V, S


def icosahedron():
    """12 vertices with the north pole first and the south pole last, 20 CCW faces."""
    z, r = 1 / np.sqrt(5.0), 2 / np.sqrt(5.0)
    a = 2 * np.pi * np.arange(5) / 5
    up = np.stack([r * np.cos(a), r * np.sin(a), np.full(5, z)], 1)
    lo = np.stack([r * np.cos(a + np.pi / 5), r * np.sin(a + np.pi / 5), np.full(5, -z)], 1)
    V = np.vstack([[0, 0, 1.0], up, lo, [0, 0, -1.0]])
    F = []
    for k in range(5):
        u0, u1, l0, l1 = 1 + k, 1 + (k + 1) % 5, 6 + k, 6 + (k + 1) % 5
        F += [[0, u0, u1], [u0, l0, u1], [u1, l0, l1], [11, l1, l0]]
    return V, orient(V, np.array(F))


class IcosahedralMesh:
    """Level-k icosahedral mesh (stripy.spherical_meshes.icosahedral_mesh's
    role): ``points`` (Nv,3) with the north pole first and the south pole
    last, ``simplices`` (Ntri,3) counter-clockwise."""

    def __init__(self, refinement_levels=0):
        V, F = icosahedron()
        V, S = refine(V, F, refinement_levels)
        # move the south pole (index 11) to the end
        n = len(V)
        perm = np.concatenate([np.arange(11), np.arange(12, n), [11]])
        inv = np.empty(n, dtype=np.int64)
        inv[perm] = np.arange(n)
        self.points = V[perm]
        self.simplices = inv[S]
        self.refinement_levels = refinement_levels


# ----------------------------------------------------------------------------
# MWSS structured mesh (closed form of mesh_functions.py:352-412)
# ----------------------------------------------------------------------------
def mwss_thetas(L):
    return np.array([2.0 * t * np.pi / (2.0 * L) for t in range(L + 1)])


def mwss_phis(L):
    return np.array([2.0 * p * np.pi / (2.0 * L) for p in range(2 * L)])


def mwss_simplices(L):
    """grid_assemble (mesh_functions.py:380-412) in closed form: (4L(L-1), 3)."""
    N, M = 2 * L, L + 1
    c = np.arange(N)
    c1 = (c + 1) % N
    R = lambda r, cc: 1 + (r - 1) * N + cc
    first = np.stack([np.zeros(N, dtype=np.int64), R(1, c), R(1, c1)], 1)
    rows = []
    for r in range(1, M - 2):
        A = np.stack([R(r, c), R(r + 1, c1), R(r, c1)], 1)
        B = np.stack([R(r, c), R(r + 1, c), R(r + 1, c1)], 1)
        rows.append(np.stack([A, B], 1).reshape(-1, 3))
    S = (M - 2) * N + 1
    last = np.stack([R(M - 2, c), np.full(N, S), R(M - 2, c1)], 1)
    return np.concatenate([first] + rows + [last]).astype(np.int64)


def mwss_msimplices(L):
    """full_assembly's msimplices (mesh_functions.py:352-378): (L, 2L, 2, 4)."""
    N, M = 2 * L, L + 1
    simps = mwss_simplices(L)
    T = len(simps)
    sp1 = np.concatenate([simps, np.arange(T)[:, None]], 1)
    ms = np.empty((M - 1, N, 2, 4), dtype=np.int64)
    ms[0, :, 0] = sp1[:N]
    ms[0, :, 1] = sp1[:N]
    mid = sp1[N:T - N].reshape(M - 3, N, 2, 4)
    ms[1:M - 2] = mid
    ms[M - 2, :, 0] = sp1[T - N:]
    ms[M - 2, :, 1] = sp1[T - N:]
    return ms
