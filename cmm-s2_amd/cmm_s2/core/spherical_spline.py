"""Powell-Sabin split-spline interpolants (cmm_s2/core/spherical_spline.py API).

Classes keep the reference's names, constructor arguments and array layouts:
jets ``vals`` (.., 4, Nv) with the transposed-Jacobian storage of vector
fields, Bernstein-Bezier ``coeffs`` (.., 19, Ntri).  Device state adds a
``packed`` evaluation layout (Ntri, 6, ncomp, 6): for every split
sub-triangle the 6 coefficients the evaluation reads, contiguous (144 bytes
for a 3-component map), so one query reads one block.  ``coeffs`` is filled
only when the caller passes a buffer (as the reference's caller-allocated
arrays) or reads the property.
"""
import numpy as np
import torch

from .. import _native as N
from . import _backend as B
from . import utils

# spherical_spline.py:660-666 (kept for callers; the kernels use the same tables)
Cs = np.array([[[0, 0, 0, 0, 0, 0], [18, 1, 13, 7, 6, 16], [13, 2, 18, 11, 10, 16]],
               [[18, 14, 0, 17, 5, 4], [0, 0, 0, 0, 0, 0], [14, 18, 2, 17, 10, 9]],
               [[0, 12, 18, 3, 15, 4], [1, 18, 12, 7, 15, 8], [0, 0, 0, 0, 0, 0]]])
edges_ps = np.array([[[0, 0, 0], [3, 1, 5], [5, 2, 3]],
                     [[3, 6, 0], [0, 0, 0], [6, 3, 2]],
                     [[0, 4, 3], [1, 3, 4], [0, 0, 0]]])


def bary_minmax(X):
    """spherical_spline.py:691-697."""
    return np.argmin(X, axis=1), np.argmax(X, axis=1)


def new_vpoints(v_pts, bmm):
    """spherical_spline.py:668-689 (host helper)."""
    v1, v2, v3 = v_pts[:, 0, :], v_pts[:, 1, :], v_pts[:, 2, :]
    nrm = lambda a: a / np.linalg.norm(a, axis=1)[:, None]
    EE = np.array([v1, v2, v3, nrm((v1 + v2 + v3) / 3), nrm(v1 / 2 + v2 / 2), nrm(v2 / 2 + v3 / 2),
                   nrm(v3 / 2 + v1 / 2)])
    Js = edges_ps[bmm[0], bmm[1]]
    r = np.arange(len(Js))
    return np.stack([EE[Js[:, 0], r], EE[Js[:, 1], r], EE[Js[:, 2], r]], 1)


def _is_mwss(mesh):
    return hasattr(mesh, "msimplices") and hasattr(mesh, "grid_points")


def _assemble(mesh, vals, ncomp, coeffs, packed):
    fn = "cmm_mwss_assemble" if _is_mwss(mesh) else "cmm_tri_assemble"
    N.call(fn, mesh._h, vals.data_ptr(), ncomp, B.ptr(coeffs), B.ptr(packed), B.stream())


def _canonical_out(coeffs, shape):
    """Device buffer for a caller-provided canonical coefficient array."""
    if coeffs is None:
        return None, None
    if isinstance(coeffs, torch.Tensor) and coeffs.is_cuda and coeffs.dtype == torch.float64 \
            and coeffs.is_contiguous() and tuple(coeffs.shape) == tuple(shape):
        return coeffs, None
    return B.empty(*shape), coeffs


def _write_back(dev, host):
    if host is None:
        return
    if isinstance(host, np.ndarray):
        host[...] = dev.detach().cpu().numpy().reshape(host.shape)
    else:
        host.copy_(dev.reshape(host.shape))


class sphere_diffeomorphism(object):
    """Spline map of the sphere (spherical_spline.py:155-337).  vals (3,4,Nv);
    coeffs: optional (3,19,Ntri) buffer written in place; packed: optional
    device buffer (Ntri,108) to reuse (spline_proj_sphere hands the old map's
    buffer to the new one, evolution_functions.py:113)."""

    def __init__(self, mesh, vals, coeffs=None, packed=None):
        self.mesh = mesh
        self.vals = B.to_dev(vals)
        self.packed = packed if packed is not None else B.empty(mesh.ntri, 108)
        self._coeffs = coeffs
        self.assemble_coefficients(coeffs)

    def assemble_coefficients(self, coeffs=None):
        dev, host = _canonical_out(coeffs, (3, 19, self.mesh.ntri))
        _assemble(self.mesh, self.vals, 3, dev, self.packed)
        if dev is not None:
            _write_back(dev, host)
            self._coeffs = coeffs if host is not None else dev
        return self._coeffs

    @property
    def coeffs(self):
        if self._coeffs is None:
            c = B.empty(3, 19, self.mesh.ntri)
            _assemble(self.mesh, self.vals, 3, c, None)
            self._coeffs = c
        return self._coeffs

    def eval(self, q, layout="N3", out=None, normalize=True):
        """Device evaluation; returns (N,3) for layout 'N3' or (3,N) for '3N'."""
        n = q.shape[0] if layout == "N3" else q.shape[1]
        if out is None:
            out = B.empty(n, 3) if layout == "N3" else B.empty(3, n)
        sp, sc = B.point_strides(q, layout)
        osp, osc = B.point_strides(out, layout)
        N.call("cmm_tri_eval", self.mesh._h, self.packed.data_ptr(), 3, int(normalize), q.data_ptr(), n,
               sp, sc, out.data_ptr(), osp, osc, B.stream())
        return out

    def __call__(self, q_pts):
        """spherical_spline.py:178-205: q (N,3) -> normalised (3,N)."""
        q = B.to_dev(q_pts)
        n = q.shape[0]
        out = B.empty(3, n)
        sp, sc = B.point_strides(q, "N3")
        N.call("cmm_tri_eval", self.mesh._h, self.packed.data_ptr(), 3, 1, q.data_ptr(), n, sp, sc,
               out.data_ptr(), 1, n, B.stream())
        return B.like(out, q_pts)

    def eval_grad(self, q_pts, eval_pts=None):
        """spherical_spline.py:208-265: differential of the map at q (N,3) ->
        (3,3,N), [c][d] = d phi_c / d x_d on the tangent plane at q (the map
        value is recomputed, ``eval_pts`` is accepted for the reference's
        signature and ignored)."""
        q = B.to_dev(q_pts)
        n = q.shape[0]
        out = B.empty(3, 3, n)
        sp, sc = B.point_strides(q, "N3")
        N.call("cmm_tri_eval_grad", self.mesh._h, self.packed.data_ptr(), 3, q.data_ptr(), n, sp, sc,
               out.data_ptr(), B.stream())
        return B.like(out, q_pts)

    def stencil_eval(self, q_pts, st_pts):
        """spherical_spline.py:268-298: locate the vertex feet q (3,Nv), evaluate
        the four stencil feet st (3,4,Nv) on that split sub-triangle's
        polynomial (not normalised); st is overwritten and returned."""
        q = B.to_dev(q_pts)
        st = B.to_dev(st_pts)
        nv = self.mesh.nv
        jets = B.empty(3, 4, nv)
        ev = B.empty(3, 4, nv)
        N.call("cmm_tri_stencil_jets", self.mesh._h, self.packed.data_ptr(), q.data_ptr(), q.stride(0),
               st.data_ptr(), st.stride(0), self.mesh.d_tan.data_ptr(), 1e-5, jets.data_ptr(),
               ev.data_ptr(), B.stream())
        if isinstance(st_pts, np.ndarray):
            st_pts[...] = ev.cpu().numpy()
            return st_pts
        if isinstance(st_pts, torch.Tensor) and st_pts.is_cuda and st_pts.dtype == torch.float64:
            st_pts.copy_(ev)
            return st_pts
        return ev


class composite_sphere_diffeomorphism(object):
    """Chain of submaps (spherical_spline.py:340-508).  vals (n,3,4,Nv);
    coeffs optional (n,3,19,Ntri); ns the remap schedule; Nc filled count."""

    def __init__(self, mesh, vals, coeffs, ns):
        self.mesh = mesh
        self.vals = B.to_dev(vals)
        self.ns = ns
        self._coeffs = coeffs
        n = self.vals.shape[0]
        self.packed = B.empty(n, mesh.ntri, 108)
        self.Nc = 0

    @property
    def coeffs(self):
        if self._coeffs is None:
            n = self.vals.shape[0]
            c = B.zeros(n, 3, 19, self.mesh.ntri)
            for i in range(self.Nc):
                _assemble(self.mesh, self.vals[i], 3, c[i], None)
            self._coeffs = c
        return self._coeffs

    def assemble_coefficients(self, i, vals):
        """spherical_spline.py:471-508."""
        dst = self.vals[i]
        if isinstance(vals, torch.Tensor) and vals.is_cuda and vals.dtype == torch.float64 and vals.is_contiguous() \
                and tuple(vals.shape) == tuple(dst.shape) and dst.is_contiguous():
            # a library kernel (it joins a member-batched step in order)
            N.call("cmm_copy_f64", dst.data_ptr(), vals.data_ptr(), dst.numel(), B.stream())
        else:
            self.vals[i] = B.to_dev(vals)
        dev = None
        if self._coeffs is not None:
            dev, host = _canonical_out(self._coeffs[i], (3, 19, self.mesh.ntri))
        _assemble(self.mesh, self.vals[i], 3, dev, self.packed[i])
        if dev is not None:
            _write_back(dev, host)

    def levels(self, N_c=None):
        """Packed device levels in application order (N_c-1 first)."""
        if N_c is None:
            N_c = self.Nc
        return [self.packed[i] for i in reversed(range(N_c))]

    def compose(self, q, layout="N3", out=None, pre=(), N_c=None, inter=None, grid=None, located=None):
        """Fused device composition: apply maps ``pre`` (packed buffers), then
        submaps N_c-1 .. 0.  ``located``: the first level's location of the
        grid points (split_locate(mesh, q)), used instead of a search."""
        lv = list(pre) + self.levels(N_c)
        return compose_packed(self.mesh, lv, q, layout, out, inter, grid, located)

    def __call__(self, q_pts, N_c=None):
        """spherical_spline.py:365-398: q (N,3) -> (N,3)."""
        q = B.to_dev(q_pts)
        return B.like(self.compose(q, "N3", N_c=N_c), q_pts)

    def jacobian(self, q_pts, N_c=None, pre=()):
        """Composed map and its differential through the whole chain:
        q (N,3) -> positions (N,3) and (3,3,N) [c][d] (tangent planes at q and
        at the image).  Its determinant on the tangent planes is the area
        distortion (1 for incompressible flows)."""
        q = B.to_dev(q_pts)
        n = q.shape[0]
        lv = list(pre) + self.levels(N_c)
        out = B.empty(n, 3)
        jac = B.empty(3, 3, n)
        sp, sc = B.point_strides(q, "N3")
        arr = N.ptr_array([p.data_ptr() for p in lv])
        N.call("cmm_tri_compose_grad", self.mesh._h, arr, len(lv), q.data_ptr(), n, sp, sc, out.data_ptr(), 3, 1,
               jac.data_ptr(), B.stream())
        return B.like(out, q_pts), B.like(jac, q_pts)

    def eval_all(self, q_pts):
        """spherical_spline.py:400-437: [q, after level Nc-1, ..., after level 0]."""
        q = B.to_dev(q_pts)
        n = q.shape[0]
        inter = B.empty(max(1, self.Nc), n, 3)
        self.compose(q, "N3", N_c=self.Nc, inter=inter)
        outs = [q.clone()] + [inter[i] for i in range(self.Nc)]
        return [B.like(o, q_pts) for o in outs]


# tile rows chosen by tune_grid_tiles per (mesh, grid shape, depth)
_tiles = {}


def split_locate(mesh, q):
    """Location of the points q (N,3) for evaluation on ``mesh``: (block
    int32 (N,) = 6 triangle + split sub-triangle, -1 if not located;
    sub-triangle barycentrics (N,3)), as the evaluation computes them."""
    q = B.to_dev(q)
    n = q.shape[0]
    sp, sc = B.point_strides(q, "N3")
    blk = torch.empty(n, dtype=torch.int32, device=q.device)
    cb = B.empty(n, 3)
    N.call("cmm_tri_split_locate", mesh._h, q.data_ptr(), n, sp, sc, blk.data_ptr(), cb.data_ptr(), B.stream())
    return blk, cb


def tune_grid_tiles(mesh, levels, q, grid, out=None, reps=3, candidates=(1, 2, 4, 8, 16), located=None):
    """Time the grid composition with each tile shape (rows x 64/rows) on
    these inputs and keep the fastest for (mesh, grid shape, depth): a
    measured choice of the launch shape, results are identical for every
    candidate.  Returns {rows: seconds}."""
    import time
    n = q.shape[0]
    if out is None:
        out = B.empty(n, 3)
    key = (mesh._h, int(grid[0]), int(grid[1]), len(levels), located is not None)
    times = {}
    for rows in candidates:
        _tiles[key] = rows
        compose_packed(mesh, levels, q, "N3", out, grid=grid, located=located)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            compose_packed(mesh, levels, q, "N3", out, grid=grid, located=located)
        torch.cuda.synchronize()
        times[rows] = (time.perf_counter() - t0) / reps
    _tiles[key] = min(times, key=times.get)
    return times


def compose_packed(mesh, levels, q, layout="N3", out=None, inter=None, grid=None, located=None):
    """Fused composition through packed levels (application order).  ``grid``
    = (nring, ncol) declares q a ring-major grid: same result, tiled
    traversal.  ``located`` (grid only): split_locate(mesh, q), the first
    level's location of the points: same result without its search."""
    n = q.shape[0] if layout == "N3" else q.shape[1]
    if out is None:
        out = B.empty(n, 3) if layout == "N3" else B.empty(3, n)
    sp, sc = B.point_strides(q, layout)
    osp, osc = B.point_strides(out, layout)
    arr = N.ptr_array([p.data_ptr() for p in levels])
    if grid is not None:
        nring, ncol = int(grid[0]), int(grid[1])
        if nring * ncol != n:
            raise N.CmmError(f"grid {grid} does not hold {n} points")
        rows = _tiles.get((mesh._h, nring, ncol, len(levels), located is not None), 4)
        if located is not None:
            blk, cb = located
            if blk.shape[0] != n or cb.shape != (n, 3) or not cb.is_contiguous():
                raise N.CmmError("located does not hold the grid's points")
            N.call("cmm_tri_compose_grid_located", mesh._h, arr, len(levels), q.data_ptr(), nring, ncol, rows, sp,
                   sc, out.data_ptr(), osp, osc, B.ptr(inter), blk.data_ptr(), cb.data_ptr(), B.stream())
        else:
            N.call("cmm_tri_compose_grid", mesh._h, arr, len(levels), q.data_ptr(), nring, ncol, rows, sp, sc,
                   out.data_ptr(), osp, osc, B.ptr(inter), B.stream())
    else:
        if located is not None:
            raise N.CmmError("located needs grid=")
        N.call("cmm_tri_compose", mesh._h, arr, len(levels), q.data_ptr(), n, sp, sc, out.data_ptr(), osp,
               osc, B.ptr(inter), B.stream())
    return out


def lagrange_weights(nVs, t, dt, ts0):
    """Time weights of spline_interp_velocity.__call__ (spherical_spline.py:596-612)."""
    if nVs == 3:
        tau0, tau1, tau2 = ts0, ts0 + dt, ts0 + 2 * dt
        return ((t - tau1) * (t - tau2) / (2 * dt ** 2), (t - tau0) * (t - tau2) / (-dt ** 2),
                (t - tau0) * (t - tau1) / (2 * dt ** 2))
    if nVs == 2:
        tau = (t - ts0) / dt
        return (1 - tau, tau, 0.0)
    return (1.0, 0.0, 0.0)


class spline_interp_velocity(object):
    """2+1-D velocity spline (spherical_spline.py:545-655) on an MWSS mesh.
    vals (3 levels, 3, 4, Nu); coeffs optional; ts the time grid.  Time levels
    live in a 3-slot ring: ``shift()`` rotates slots instead of copying the
    1.4 GB of level data (spherical_harmonic_tools.py:73-76 copies).

    ``sparse = True`` (the CMM loop sets it): a level's packed blocks are
    written only on a set of triangles and the rows next to the poles
    (``cmm_mwss_assemble_sparse``), with a validity bit per slot and
    triangle; an evaluation elsewhere builds its block from the level's jets
    with the same arithmetic, so results do not depend on the set.  With
    ``mark_every`` the set is the triangles the evaluations since the
    previous level located; otherwise it is the previous level's set plus
    the triangles whose block some evaluation missed.  The feet of a fixed
    set of points move little from one step to the next, so most evaluations
    find their block assembled."""

    def __init__(self, mesh, vals, coeffs, ts):
        self.mesh = mesh
        self._vals = B.to_dev(vals)
        self._packed = B.zeros(3, mesh.ntri, 108)
        self._order = [0, 1, 2]
        self._coeffs_buf = coeffs
        self.ts = np.array(ts, dtype=np.float64)
        self.stepping = False
        self.nVs = 1
        self.sparse = False
        self._vbits = torch.full((mesh.ntri,), 7, dtype=torch.uint8, device=self._packed.device)
        self._mark = torch.ones(mesh.ntri, dtype=torch.uint8, device=self._packed.device)
        self.mark_every = True
        self._newest = -1  # physical slot of the last sparse assembly

    def _assemble_sparse(self, i, mark):
        inherit = -1 if (mark is None or self.mark_every or self._newest == self._order[i]) else self._newest
        N.call("cmm_mwss_assemble_sparse", self.mesh._h, self.level_vals(i).data_ptr(), self.level_packed(i).data_ptr(),
               self._vbits.data_ptr(), self._order[i], inherit, B.ptr(mark), B.stream())
        if mark is not None:
            self._newest = self._order[i]

    # logical level access
    def level_vals(self, i):
        return self._vals[self._order[i]]

    def level_packed(self, i):
        return self._packed[self._order[i]]

    def level_slot(self, i):
        return self._order[i]

    def valid_fraction(self, i):
        """Fraction of triangles whose packed blocks of level i are current."""
        return int(((self._vbits >> self._order[i]) & 1).sum()) / self.mesh.ntri

    def mark_all(self):
        """Make the next sparse assembly write every triangle and the
        evaluations mark every triangle they locate (a torch fill kernel: no
        memset, so it may run under stream capture)."""
        self._mark.fill_(1)
        self.mark_every = True
        self._newest = -1

    def complete(self):
        """Assemble the triangles every level still lacks (before callers
        that read the packed blocks directly)."""
        if not self.sparse:
            return
        for i in range(3):
            self._assemble_sparse(i, None)

    def shift(self):
        """Logical levels (0,1,2) <- (1,2,stale); level 2 is rewritten next."""
        self._order = self._order[1:] + self._order[:1]

    @property
    def vals(self):
        return self._vals[self._order]

    @property
    def coeffs(self):
        c = B.empty(3, 3, 19, self.mesh.ntri)
        for i in range(3):
            _assemble(self.mesh, self.level_vals(i), 3, c[i], None)
        return c

    def init_coeffs(self, i, complete=False):
        """spherical_spline.py:565-568.  With ``sparse`` (and not
        ``complete``) only the triangles marked since the last call."""
        dev = None
        if self._coeffs_buf is not None:
            dev, host = _canonical_out(self._coeffs_buf[i], (3, 19, self.mesh.ntri))
        if self.sparse and dev is None and not complete:
            self._assemble_sparse(i, self._mark)
            return
        _assemble(self.mesh, self.level_vals(i), 3, dev, self.level_packed(i))
        self._vbits.bitwise_or_(1 << self._order[i])
        if dev is not None:
            _write_back(dev, host)

    def weights(self, t, dt):
        return lagrange_weights(self.nVs, t, dt, float(self.ts[0]))

    def levels_used(self):
        return min(self.nVs, 3) if self.nVs in (1, 2, 3) else 3

    def eval(self, t, dt, q, layout="3N", out=None):
        self.complete()
        n = q.shape[0] if layout == "N3" else q.shape[1]
        if out is None:
            out = B.empty(n, 3) if layout == "N3" else B.empty(3, n)
        nl = self.levels_used()
        w = self.weights(t, dt)
        arr = N.ptr_array([self.level_packed(i).data_ptr() for i in range(nl)])
        warr = N.dbl_array(w[:nl])
        sp, sc = B.point_strides(q, layout)
        osp, osc = B.point_strides(out, layout)
        N.call("cmm_mwss_eval", self.mesh._h, arr, warr, nl, 3, q.data_ptr(), n, sp, sc, out.data_ptr(),
               osp, osc, B.stream())
        return out

    def __call__(self, t, dt, q_pts):
        """spherical_spline.py:570-615: q (3,N) -> (3,N) (not normalised)."""
        q = B.to_dev(q_pts)
        return B.like(self.eval(t, dt, q, "3N"), q_pts)


class spherical_spline(object):
    """Scalar PS spline (spherical_spline.py:19-152).  vals (4,Nv): value and
    Cartesian gradient; works on an MWSS or an unstructured mesh."""

    def __init__(self, mesh, vals, coeffs=None):
        self.mesh = mesh
        self.vals = B.to_dev(vals)
        self.packed = B.empty(mesh.ntri, 36)
        self._coeffs = coeffs
        dev, host = _canonical_out(coeffs, (19, mesh.ntri))
        _assemble(mesh, self.vals, 1, dev, self.packed)
        if dev is not None:
            _write_back(dev, host)
        self.vort = False

    @property
    def coeffs(self):
        if self._coeffs is None:
            c = B.empty(19, self.mesh.ntri)
            _assemble(self.mesh, self.vals, 1, c, None)
            self._coeffs = c
        return self._coeffs

    def eval_xyz(self, q, layout="N3"):
        n = q.shape[0] if layout == "N3" else q.shape[1]
        out = B.empty(n)
        sp, sc = B.point_strides(q, layout)
        if _is_mwss(self.mesh):
            arr = N.ptr_array([self.packed.data_ptr()])
            N.call("cmm_mwss_eval", self.mesh._h, arr, N.dbl_array([1.0]), 1, 1, q.data_ptr(), n, sp, sc,
                   out.data_ptr(), 1, 1, B.stream())
        else:
            N.call("cmm_tri_eval", self.mesh._h, self.packed.data_ptr(), 1, 0, q.data_ptr(), n, sp, sc,
                   out.data_ptr(), 1, 1, B.stream())
        return out

    def gradient(self, q_pts):
        """spherical_spline.py:73-111: tangential gradient at q (N,3) -> (3,N)."""
        q = B.to_dev(q_pts)
        n = q.shape[0]
        out = B.empty(3, n)
        sp, sc = B.point_strides(q, "N3")
        if _is_mwss(self.mesh):
            N.call("cmm_mwss_eval_grad", self.mesh._h, self.packed.data_ptr(), q.data_ptr(), n, sp, sc,
                   out.data_ptr(), B.stream())
        else:
            N.call("cmm_tri_eval_grad", self.mesh._h, self.packed.data_ptr(), 1, q.data_ptr(), n, sp, sc,
                   out.data_ptr(), B.stream())
        return B.like(out, q_pts)

    def __call__(self, phi, the):
        """spherical_spline.py:42-70: 2-D (phi, theta) arrays -> 2-D values."""
        p, t = B.to_dev(phi), B.to_dev(the)
        shape = p.shape
        q = torch.stack(utils.sphere2cart(p.reshape(-1), t.reshape(-1)), 1).contiguous()
        return B.like(self.eval_xyz(q).reshape(shape), phi)
