"""Barotropic vorticity CMM driver (cmm_s2/core/spherical_simulations.py API).

``euler_simulation_sphere_remapping`` (spherical_simulations.py:144-297)
keeps its signature and returns the composite submap object.  The loop is
device resident: per step one fused map update (RK4 integration of vertex and
stencil feet, stencil evaluation + jet projection, assembly), one fused
composition of the current map and all submaps on the MWSS sample grid, the
vorticity sampling, one forward transform, the spectral operators, one
batched 12-field inverse transform and the velocity assembly -- no host round
trip.  ``BarotropicCMM`` exposes the same loop step by step (benchmarks).
"""
import numpy as np
import torch

from .. import _native as N
from . import _backend as B
from . import evolution_functions as evol
from .mesh_functions import structure_spherical_triangulation
from .spherical_harmonic_tools import SHT, nlm, project_stream_lm
from .spherical_spline import (composite_sphere_diffeomorphism, sphere_diffeomorphism,
                               spline_interp_velocity, spherical_spline)
from . import utils


def identity_jets(mesh):
    """spherical_simulations.py:165-169 on the device."""
    v = mesh.d_vertices.t()
    jets = B.empty(3, 4, mesh.nv)
    jets[:, 0, :] = v
    jets[:, 1:, :] = torch.stack([utils.grad_x(v), utils.grad_y(v), utils.grad_z(v)])
    return jets


class VorticitySampler(object):
    """w0(X) + 2 rot (cos theta(X) - cos theta_grid) on the device
    (spherical_simulations.py:268-272)."""

    def __init__(self, vorticity, grid_mesh, rot_rate):
        self.grid = grid_mesh
        self.rot = float(rot_rate)
        if isinstance(vorticity, spherical_spline):
            if not hasattr(vorticity.mesh, "msimplices"):
                raise N.CmmError("a spline vorticity must live on an MWSS mesh")
            self.kind, self.wmesh, self.wpacked = N.VORT_SPLINE, vorticity.mesh._h, vorticity.packed
            self.params = [0.0] * 4
            self._keep = vorticity
        elif hasattr(vorticity, "_cmm_kind") and hasattr(vorticity, "sample_xyz"):
            self.kind, self.wmesh = vorticity._cmm_kind, None
            self.wpacked = getattr(vorticity, "_cmm_table", None)  # series kinds: coefficient table
            self.params = list(vorticity._cmm_params)
            self._keep = vorticity
        elif callable(vorticity):
            # a caller's own w(phi, theta) (dynamical_fields.py:176-214 style)
            self.kind, self.wmesh, self.wpacked, self.params = None, None, None, [0.0] * 4
            self._fn = vorticity
        else:
            raise N.CmmError("vorticity must be a spherical_spline on an MWSS mesh, a "
                             "cmm_s2.core.dynamical_fields vorticity or a callable w(phi, theta)")

    def __call__(self, X, out=None, rot=None, cos_grid=None, ring_len=None):
        n = X.shape[0]
        if out is None:
            out = B.empty(n)
        cg = self.grid.d_cos_theta if cos_grid is None else cos_grid
        rl = 2 * self.grid.L if ring_len is None else ring_len
        rot = float(self.rot if rot is None else rot)
        if self.kind is None:
            return self._call_plugin(X, out, rot, cg, int(rl))
        N.call("cmm_sample_vorticity", self.kind, self.wmesh, B.ptr(self.wpacked),
               N.dbl_array(self.params), X.data_ptr(), n, X.stride(0), X.stride(1), cg.data_ptr(),
               int(rl), rot, out.data_ptr(), B.stream())
        return out

    def _call_plugin(self, X, out, rot, cos_grid, ring_len):
        """w(phi, theta) + 2 rot (cos theta(X) - cos theta_ring) for a caller's
        callable: phi, theta of X on the device (utils.cart2sphere), w called
        on device tensors, or on host arrays when it only takes NumPy (the
        reference's fields use np.* functions)."""
        from . import utils
        phi, theta = utils.cart2sphere(X.t())
        w = B.call_plugin(self._fn, phi, theta)
        ring = torch.arange(X.shape[0], device=X.device) // ring_len
        out.copy_(w.reshape(-1) + 2 * rot * torch.cos(theta) - 2 * rot * cos_grid[ring])
        return out


class BarotropicCMM(object):
    """State of euler_simulation_sphere_remapping (spherical_simulations.py:144-297)."""

    RESEED = 16

    def __init__(self, L, Nt, T, mesh, vorticity, rot_rate=2 * np.pi, n_maps=10, legendre="fp64",
                 sparse_velocity=True):
        self.L, self.Nt, self.T, self.mesh = L, Nt, T, mesh
        self.legendre = legendre
        self.rot = rot_rate
        self.ns = [i for i in range(1, Nt) if i % n_maps == 0] + [Nt]
        jet0 = identity_jets(mesh)
        vals = B.zeros(len(self.ns), 3, 4, mesh.nv)
        vals[0] = jet0
        self.map0 = sphere_diffeomorphism(mesh, jet0.clone())
        self.submaps = composite_sphere_diffeomorphism(mesh, vals, None, self.ns)
        self.tspan = np.linspace(0, T, Nt, endpoint=False)
        self.dt = self.tspan[1] - self.tspan[0]
        self.mesh_u = structure_spherical_triangulation(L)
        self.U = spline_interp_velocity(self.mesh_u, B.zeros(3, 3, 4, self.mesh_u.nv), None,
                                        ts=self.tspan[0:3])
        # velocity levels assembled where the map's feet fall (the same bits
        # as full levels; see spline_interp_velocity)
        self.U.sparse = bool(sparse_velocity)
        self.plan = SHT.get(L, legendre)
        self.sampler = VorticitySampler(vorticity, self.mesh_u, rot_rate)
        self.samples = self.mesh_u.sample_points()
        n = self.samples.shape[0]
        # the sample grid's location in the map mesh, for the first level of
        # every step's composition (the current map; mesh and grid are
        # fixed): computed here, outside any capture or batch recording
        from .spherical_spline import split_locate
        self._loc0 = split_locate(mesh, self.samples)
        self.X = B.empty(n, 3)
        self.w = B.empty(1, L + 1, 2 * L)
        self.w_lm = B.empty(nlm(L), 1, 2)
        self.ops = B.empty(nlm(L), 12, 2)
        self.fields = B.empty(12, L + 1, 2 * L)
        self.curr = self.map0
        self.identity = True
        self.count = 0

    def _velocity_from(self, X, level, rot=None):
        self.sampler(X, out=self.w.reshape(-1), rot=rot)
        self.plan.forward(self.w, out=self.w_lm)
        project_stream_lm(self.plan, self.w_lm, self.U, level, apply_dinv=True, fields=self.fields,
                          ops=self.ops)

    def reset(self):
        """Return the host-side loop state to that of a fresh object (step
        count, current map, velocity time levels, submap count).  Device
        buffers keep their contents: a run writes every value before it reads
        it, so a run after reset() gives the bits of a first run."""
        self.curr = self.map0
        self.identity = True
        self.count = 0
        self.submaps.Nc = 0
        U = self.U
        U._order = [0, 1, 2]
        U.ts = np.array(self.tspan[0:3], dtype=np.float64)
        U.stepping = False
        U.nVs = 1

    def bootstrap(self, tune=True):
        """spherical_simulations.py:216-261.  ``tune`` times the composition
        tile shapes (host synchronisation; pass False under stream capture)."""
        U, dt = self.U, self.dt
        U.mark_all()
        self._velocity_from(self.samples, 0, rot=0.0)
        int0 = evol.advect_project_sphere(self.map0, evol.Euler_step_proj, 0, dt, U, identity=True)
        int0.eval(self.samples, "N3", out=self.X)
        self._velocity_from(self.X, 1)
        U.nVs = 2
        int1 = evol.advect_project_sphere(self.map0, evol.improved_Euler_proj, 0, dt, U, identity=True)
        int1.eval(self.samples, "N3", out=self.X)
        self._velocity_from(self.X, 1)
        int2 = evol.advect_project_sphere(int1, evol.RK4_proj, dt, dt, U)
        int2.eval(self.samples, "N3", out=self.X)
        U.nVs = 3
        self._velocity_from(self.X, 2)
        self.curr = int2
        U.stepping = True
        self.identity = False
        self.count = 2
        if tune:
            self.tune_tiles()

    def tune_tiles(self):
        """Pick the tile shape of the per-step sample-grid composition for
        every depth the run will use (1 + remaps), timing the candidates on
        the current map; results do not depend on the choice."""
        from .spherical_spline import tune_grid_tiles
        grid = (self.L + 1, 2 * self.L)
        for depth in range(1, len(self.ns) + 2):
            tune_grid_tiles(self.mesh, [self.curr.packed] * depth, self.samples, grid, out=self.X, reps=2,
                            located=self.grid_location())

    def grid_location(self):
        """Location of the sample grid in the map mesh (the evaluation's own
        block indices and sub-triangle barycentrics, stored at construction)."""
        return self._loc0

    def step(self):
        """One loop iteration (spherical_simulations.py:263-294).  Returns True
        when this step ended with a remap."""
        t = self.tspan[self.count]
        # sparse velocity levels: every RESEED-th step rebuilds the assembled
        # set from all of the step's evaluations; in between it grows by the
        # triangles some evaluation missed
        self.U.mark_every = self.count % self.RESEED == 0
        self.curr = evol.advect_project_sphere(self.curr, evol.RK4_proj, t, self.dt, self.U,
                                               identity=self.identity)
        self.submaps.compose(self.samples, "N3", out=self.X, pre=[self.curr.packed], grid=(self.L + 1, 2 * self.L),
                             located=self.grid_location())
        self._velocity_from(self.X, 2)
        self.U.ts = self.U.ts + self.dt
        self.count += 1
        self.identity = False
        if self.count in self.ns:
            idx = self.ns.index(self.count)
            self.submaps.Nc = idx + 1
            self.submaps.assemble_coefficients(idx, self.curr.vals)
            self.curr = self.map0
            self.identity = True
            return True
        return False

    def run(self, checkpoint=None, every=None):
        """Bootstrap (unless resumed) and step to Nt.  With ``checkpoint`` (a
        path) and ``every`` = k, the state is saved after every k-th step and
        at the end (replaces the pickles of spherical_simulations.py:434-438,
        453-454)."""
        if self.count == 0:
            self.bootstrap()
        while self.count < self.Nt:
            self.step()
            if checkpoint is not None and every and self.count % every == 0:
                self.save(checkpoint)
        if checkpoint is not None:
            self.save(checkpoint)
        return self.submaps

    # ---- checkpoint / resume -------------------------------------------
    def save(self, path):
        """Write the loop state to a safetensors file: submap jets, current
        map jets, the three velocity time levels (logical order), counters
        and times, and the initial vorticity (spline jets, or the analytic
        field's kind and parameters).  Coefficients are not stored: they are
        re-assembled on resume, which reproduces them bit for bit."""
        from . import checkpoint as ck
        ck.save_state(self, path)

    @classmethod
    def resume(cls, path, mesh, vorticity=None):
        """Rebuild a BarotropicCMM from ``save`` on the icosahedral ``mesh``
        (the map mesh is an input, as in the reference).  ``vorticity`` may be
        omitted when the file holds it."""
        from . import checkpoint as ck
        return ck.load_state(cls, path, mesh, vorticity)

    def zoom_vorticity(self, x0=3.22055, y0=1.1963, ks=(-2, -4, -7, -10, -13), n=1500, same_center=True):
        """Absolute vorticity w0(X) + 2 rot cos theta(X) of the composed map on
        zoom windows (figures/zoom_sampling_figure.py:63-90): for each k an
        n x n (phi, theta) grid of half-width 2^k, phi = linspace(x0 - 2^k,
        x0 + 2^k, n) along columns.  The reference centres theta on x0 as
        well (its `ys` repeats `x0`); ``same_center=False`` centres it on y0.
        Returns a device tensor (len(ks), n, n) [window, theta row, phi col]."""
        nw = len(ks)
        pts = []
        for k in ks:
            h = 2.0 ** k
            xs = np.linspace(x0 - h, x0 + h, n, endpoint=True)
            yc = x0 if same_center else y0
            ys = np.linspace(yc - h, yc + h, n, endpoint=True)
            PH, TH = np.meshgrid(xs, ys)
            pts.append(np.stack([np.sin(TH) * np.cos(PH), np.sin(TH) * np.sin(PH), np.cos(TH)], -1).reshape(-1, 3))
        q = B.to_dev(np.concatenate(pts))
        X = self.submaps.compose(q, "N3", pre=[] if self.identity else [self.curr.packed])
        cos_q = q[:, 2].contiguous()  # cos theta of each evaluation point, one point per "ring"
        w = self.sampler(X, cos_grid=cos_q, ring_len=1)  # w0(X) + 2 rot (cos theta(X) - cos theta(q))
        return (w + 2 * self.rot * cos_q).reshape(nw, n, n)

    def conservation_errors(self, L_out, grid=None, true_vorticity=None):
        """Relative enstrophy and energy errors of the absolute vorticity
        w0(X) + 2 rot cos theta(X) on the MWSS grid of bandlimit L_out
        against the same quantities of w0 + 2 rot cos theta at t = 0, and,
        with ``true_vorticity`` (the exact relative vorticity on that grid),
        the relative max-norm error: figures/data_gen_convergence.py:58-95."""
        from .spherical_harmonic_tools import enstrophy_energy
        g = grid if grid is not None else structure_spherical_triangulation(L_out)
        q = g.sample_points()
        cos = g.d_cos_theta[:, None]
        w0 = self.sampler(q, cos_grid=g.d_cos_theta, ring_len=2 * L_out).reshape(L_out + 1, 2 * L_out)
        wT = self.vorticity_on_grid(L_out, grid=g)
        e0, n0 = enstrophy_energy(w0 + 2 * self.rot * cos, L_out)
        eT, nT = enstrophy_energy(wT + 2 * self.rot * cos, L_out)
        out = {"enstrophy_rel_err": abs(eT - e0) / e0, "energy_rel_err": abs(nT - n0) / n0,
               "enstrophy": eT, "energy": nT, "enstrophy_0": e0, "energy_0": n0}
        if true_vorticity is not None:
            t = B.to_dev(true_vorticity).reshape(L_out + 1, 2 * L_out)
            out["max_rel_err"] = float((wT - t).abs().max() / t.abs().max())
        return out

    def vorticity_on_grid(self, L_out, out=None, grid=None):
        """Relative vorticity w0(X) + 2 rot (cos theta(X) - cos theta) of the
        composed map on the MWSS grid of bandlimit L_out (the upsampling of
        figures/data_gen_experiments.py:122-144)."""
        g = grid if grid is not None else structure_spherical_triangulation(L_out)
        q = g.sample_points()
        X = self.submaps.compose(q, "N3", pre=[] if self.identity else [self.curr.packed],
                                 grid=(L_out + 1, 2 * L_out))
        return self.sampler(X, out=out, cos_grid=g.d_cos_theta, ring_len=2 * L_out).reshape(L_out + 1, 2 * L_out)


def euler_simulation_sphere_remapping(L, Nt, T, mesh, vorticity, rot_rate=2 * np.pi, n_maps=10):
    """spherical_simulations.py:144-297: returns the composite submap object."""
    sim = BarotropicCMM(L, Nt, T, mesh, vorticity, rot_rate, n_maps)
    return sim.run()
