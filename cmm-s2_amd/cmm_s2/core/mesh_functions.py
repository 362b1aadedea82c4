"""Meshes and point location (cmm_s2/core/mesh_functions.py API).

``spherical_triangulation`` (mesh_functions.py:15-166) and
``structure_spherical_triangulation`` (169-318) keep the reference's
constructor arguments, attributes and ``query`` contract.  Set-up (hull,
refinement, stencil points) runs on the host once; every ``query`` runs on
the GPU (libcmm_s2_amd.so): a cube-map bucket search with exact barycentric
tests for unstructured meshes, the closed-form numpy-faithful cell search for
the MWSS mesh.
"""
import numpy as np
import torch

from .. import _native as N
from . import _backend as B
from . import _meshgen as G
from . import utils


def icosahedral_mesh(refinement_levels=0):
    """Icosahedral point set and triangles (north pole first, south pole last)."""
    return G.IcosahedralMesh(refinement_levels)


def sphere_tan_proj(a, b):
    """mesh_functions.py:320-329: unit tangential part of a (N,3) at b (N,3) -> (3,N)."""
    c = utils.dot(b.T, a.T)
    d = np.sqrt(utils.dot(a.T, a.T) - 2 * c ** 2 + utils.dot(b.T, b.T) * c ** 2)
    return np.array([(a[:, 0] - c * b[:, 0]) / d, (a[:, 1] - c * b[:, 1]) / d, (a[:, 2] - c * b[:, 2]) / d])


def ps_split_tables(vertices, simplices):
    """mesh_functions.py:92-126: the Powell-Sabin geometry tables (Hs, Gs, Es).
    Host arrays for callers that read them; the GPU assembly recomputes the
    same quantities per triangle instead of reading 528 bytes of tables."""
    v = vertices[np.asarray(simplices)]
    v1, v2, v3 = v[:, 0].copy(), v[:, 1].copy(), v[:, 2].copy()
    v4 = utils.div_norm((v1 + v2 + v3) / 3).T
    e1, e2, e3 = (utils.div_norm(v1 / 2 + v2 / 2).T, utils.div_norm(v2 / 2 + v3 / 2).T,
                  utils.div_norm(v3 / 2 + v1 / 2).T)
    stp = sphere_tan_proj
    h12, h21, h23 = stp(v2 - v1, v1).T, stp(v1 - v2, v2).T, stp(v3 - v2, v2).T
    h32, h13, h31 = stp(v2 - v3, v3).T, stp(v3 - v1, v1).T, stp(v1 - v3, v3).T
    h41, h42, h43 = stp(v4 - v1, v1).T, stp(v4 - v2, v2).T, stp(v4 - v3, v3).T
    bc = utils.bary_coords
    Gs = [bc(v1, e1, v4, h12), bc(v2, v4, e1, h21), bc(v2, e2, v4, h23), bc(v3, v4, e2, h32),
          bc(v1, v4, e3, h13), bc(v3, e3, v4, h31), bc(v1, v4, e3, stp(v4, v1).T),
          bc(v2, v4, e1, stp(v4, v2).T), bc(v3, v4, e2, stp(v4, v3).T), bc(v1, v2, v3, v4)]
    Hs = [h12.T, h21.T, h23.T, h32.T, h13.T, h31.T, h41.T, h42.T, h43.T]
    Es = [bc(v1, v2, v3, e1), bc(v1, v2, v3, e2), bc(v1, v2, v3, e3)]
    return Hs, Gs, Es


def stencil_points(verts, eps=1e-5):
    """mesh_functions.py:128-166: epsilon-stencil points (3,4,Nv) and the
    tangent frames [a1_n, a2_n] (the pole rows 0 and -1 get the reference's
    fixed frames)."""
    a1 = np.array(utils.cross(verts.T, [0, 0, 1])).T
    a2 = np.array(utils.cross(verts.T, a1.T)).T
    a1[0, :] = utils.cross(verts[0, :], [1 / np.sqrt(2), 0, 1 / np.sqrt(2)])
    a1[-1, :] = utils.cross(verts[-1, :], [1 / np.sqrt(2), 0, -1 / np.sqrt(2)])
    a2[0, :] = utils.cross(verts[0, :], a1[0, :])
    a2[-1, :] = utils.cross(verts[-1, :], a1[-1, :])
    a1_n, a2_n = utils.div_norm(a1).T, utils.div_norm(a2).T
    e1p = verts - a1_n * eps - a2_n * eps
    e1m = verts + a1_n * eps - a2_n * eps
    e2p = verts - a1_n * eps + a2_n * eps
    e2m = verts + a1_n * eps + a2_n * eps
    sp = [utils.pi_proj(e, verts) for e in (e1p, e1m, e2p, e2m)]
    s_pts = np.array([np.array([sp[0][c], sp[1][c], sp[2][c], sp[3][c]]) for c in range(3)])
    return s_pts, [a1_n, a2_n]


class spherical_triangulation(object):
    """Spherical triangulation of points ``pts`` (N,3) with optional uniform
    midpoint refinement (mesh_functions.py:15-66).  ``simplices`` may be
    given to skip the triangulation (e.g. from ``icosahedral_mesh``).
    Otherwise ``triangulation`` picks how it is built: "hull" (default) takes
    the convex hull (seconds at any size; the same triangles as the
    reference, in Qhull's hull order), "delaunay" the reference's own call,
    Qhull's Delaunay tessellation of the origin and the points with its row
    order (mesh_functions.py:41-56; ``qhull_options`` as scipy takes them,
    None = the reference's defaults, which take minutes from ~40,000 points
    on), so that ``query()[1]`` equals the reference's triangle indices; its
    Delaunay object is kept as ``mesh_d``."""

    def __init__(self, pts, refinement_levels=0, simplices=None, cells_per_face=0, triangulation="hull",
                 qhull_options=None):
        pts = np.asarray(pts, dtype=np.float64)
        mesh_d = None
        if simplices is not None:
            simplices = G.orient(pts, simplices)
        elif triangulation == "delaunay":
            simplices, mesh_d = G.delaunay_triangles(pts, qhull_options)
        elif triangulation == "hull":
            simplices = G.hull_triangles(pts)
        else:
            raise N.CmmError('triangulation must be "hull" or "delaunay"')
        V, S = G.refine(pts, simplices, refinement_levels)
        self.vertices = np.ascontiguousarray(V)
        self.simplices = np.ascontiguousarray(S, dtype=np.int64)
        self.x, self.y, self.z = V[:, 0], V[:, 1], V[:, 2]
        self.mesh_d = mesh_d if refinement_levels == 0 else None
        self._ps = None
        self.stencil_pts()
        h = N.vp()
        N.call("cmm_tri_mesh_create", N.dptr(self.vertices), len(V), N.lptr(self.simplices),
               len(S), int(cells_per_face), C_byref(h))
        self._h = h.value
        dev = B.device()
        self.d_vertices = torch.from_numpy(self.vertices).to(dev)
        self.d_verts3 = self.d_vertices.t().contiguous()                  # (3, Nv)
        self.d_spts = torch.from_numpy(self.s_pts.reshape(3, -1)).to(dev).contiguous()  # (3, 4Nv)
        self.d_tan = torch.from_numpy(np.stack(self.tan_vects)).to(dev).contiguous()  # (2, Nv, 3)
        self.d_simplices = torch.from_numpy(self.simplices).to(dev)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            try:
                N.lib().cmm_tri_mesh_destroy(h)
            except Exception:
                pass
            self._h = None

    @property
    def nv(self):
        return len(self.vertices)

    @property
    def ntri(self):
        return len(self.simplices)

    @property
    def ps_split(self):
        if self._ps is None:
            self._ps = ps_split_tables(self.vertices, self.simplices)
        return self._ps

    def stats(self):
        out = np.zeros(6, dtype=np.int64)
        N.call("cmm_tri_mesh_stats", self._h, N.lptr(out))
        keys = ["nv", "ntri", "cells_per_face", "total_candidates", "max_candidates", "device_bytes"]
        return dict(zip(keys, out.tolist()))

    def stencil_pts(self, eps=1e-5):
        self.s_pts, self.tan_vects = stencil_points(self.vertices, eps)

    def locate(self, q, layout="N3"):
        """Device location: (tri int32 (N,), bcc (N,3)) for device points."""
        n = q.shape[0] if layout == "N3" else q.shape[1]
        tri = B.empty(n, dtype=torch.int32)
        bcc = B.empty(n, 3)
        sp, sc = B.point_strides(q, layout)
        N.call("cmm_tri_locate", self._h, q.data_ptr(), n, sp, sc, tri.data_ptr(), bcc.data_ptr(), B.stream())
        return tri, bcc

    def query(self, q_pts):
        """mesh_functions.py:68-88: (bcc (N,3), trangs (N,), vs (N,3,3))."""
        q = B.to_dev(q_pts)
        tri, bcc = self.locate(q, "N3")
        t64 = tri.long()
        vs = self.d_vertices[self.d_simplices[t64.clamp(min=0)]]
        return B.like(bcc, q_pts), B.like(t64, q_pts), B.like(vs, q_pts)


def C_byref(h):
    import ctypes
    return ctypes.byref(h)


class structure_spherical_triangulation(object):
    """MWSS velocity mesh of bandlimit L (mesh_functions.py:169-194): vertices
    [N pole, rings theta_1..theta_{L-1} x phi_0..phi_{2L-1}, S pole]."""

    def __init__(self, L):
        self.L = L
        thetas, phis = G.mwss_thetas(L), G.mwss_phis(L)
        self.theta, self.phi = thetas, phis
        Phi, The = np.meshgrid(phis, thetas)
        self.grid_points = utils.sphere2cart(Phi, The)
        P2, T2 = np.meshgrid(phis, thetas[1:-1])
        ring = np.array(utils.sphere2cart(P2.reshape(-1), T2.reshape(-1))).T
        npole = np.array(utils.sphere2cart(np.zeros(1), thetas[:1])).T
        spole = np.array(utils.sphere2cart(np.zeros(1), thetas[-1:])).T
        self.vertices = np.vstack([npole, ring, spole])
        self._simplices = None
        self._msimplices = None
        self._ps = None
        self.sin_t, self.cos_t = np.sin(thetas), np.cos(thetas)
        self.sin_p, self.cos_p = np.sin(phis), np.cos(phis)
        h = N.vp()
        N.call("cmm_mwss_mesh_create", int(L), N.dptr(thetas), N.dptr(phis), N.dptr(self.sin_t),
               N.dptr(self.cos_t), N.dptr(self.sin_p), N.dptr(self.cos_p), C_byref(h))
        self._h = h.value
        self.d_cos_theta = B.to_dev(self.cos_t)
        self._d_samples = None

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            try:
                N.lib().cmm_mwss_mesh_destroy(h)
            except Exception:
                pass
            self._h = None

    @property
    def nv(self):
        return 2 * self.L * (self.L - 1) + 2

    @property
    def ntri(self):
        return 4 * self.L * (self.L - 1)

    @property
    def simplices(self):
        if self._simplices is None:
            self._simplices = G.mwss_simplices(self.L)
        return self._simplices

    @property
    def msimplices(self):
        if self._msimplices is None:
            self._msimplices = G.mwss_msimplices(self.L)
        return self._msimplices

    @property
    def ps_split(self):
        if self._ps is None:
            self._ps = ps_split_tables(self.vertices, self.simplices)
        return self._ps

    def sample_points(self):
        """(N,3) device MWSS sample grid, row-major [theta, phi]
        (spherical_simulations.py:210-214)."""
        if self._d_samples is None:
            g = np.stack([self.grid_points[c].reshape(-1) for c in range(3)], 1)
            self._d_samples = B.to_dev(g)
        return self._d_samples

    def locate(self, q, layout="N3"):
        n = q.shape[0] if layout == "N3" else q.shape[1]
        tri = B.empty(n, dtype=torch.int32)
        bcc = B.empty(n, 3)
        sp, sc = B.point_strides(q, layout)
        N.call("cmm_mwss_locate", self._h, q.data_ptr(), n, sp, sc, tri.data_ptr(), bcc.data_ptr(), B.stream())
        return tri, bcc

    def query(self, q_pts):
        """mesh_functions.py:196-282."""
        q = B.to_dev(q_pts)
        tri, bcc = self.locate(q, "N3")
        t64 = tri.long()
        d_simp = torch.from_numpy(self.simplices).to(q.device)
        d_vert = torch.from_numpy(self.vertices).to(q.device)
        vs = d_vert[d_simp[t64.clamp(min=0)]]
        return B.like(bcc, q_pts), B.like(t64, q_pts), B.like(vs, q_pts)


def grid_assemble(N_, M):
    """mesh_functions.py:380-412 for the MWSS layout N_ = 2L, M = L+1."""
    return G.mwss_simplices(M - 1).tolist()


def full_assembly(N_, M):
    """mesh_functions.py:352-378."""
    L = M - 1
    return G.mwss_simplices(L).tolist(), G.mwss_msimplices(L)
