"""Barotropic vorticity CMM driver with remapping (oracle).
Restates euler_simulation_sphere_remapping (spherical_simulations.py:144-297).

Pinned by tests/test_oracle_driver.py: the Rossby-Haurwitz (5,4) wave is an
exact travelling solution of the barotropic vorticity equation at angular
speed 2 Omega / 30 (dynamical_fields.py:205-210; reference
tests/test_barotropic_vorticity.py:40-43) -- the oracle's vorticity after a
short run matches it to the discretisation error and converges with mesh
refinement; a static RH wave with rot_rate 0 stays put.
"""
import numpy as np

from . import geometry as g
from . import evolution as ev
from . import meshes
from . import sht
from . import splines as sp


class Composite:
    """composite_sphere_diffeomorphism (spherical_spline.py:340-508)."""

    def __init__(self, mesh, vals, ns):
        self.mesh = mesh
        self.vals = vals
        self.ns = ns
        self.coeffs = [None] * len(ns)
        self.Nc = 0

    def assemble_coefficients(self, i, vals):
        self.vals[i] = vals
        self.coeffs[i] = sp.assemble(vals, self.mesh.simplices, self.mesh.ps_split)

    def __call__(self, q_pts, N_c=None):
        if N_c is None:
            N_c = self.Nc
        return sp.compose_eval(self.mesh.query, self.coeffs, q_pts, N_c)


def euler_simulation_sphere_remapping(L, Nt, T, mesh, vorticity, rot_rate=2 * np.pi,
                                      n_maps=10, step_hook=None, log=None):
    """spherical_simulations.py:144-297.  ``step_hook(count, curr_map, U)`` is
    an optional test probe called after every loop step; ``log(msg)`` an
    optional progress callback for the set-up and bootstrap stages."""
    log = log or (lambda msg: None)
    ns = [i for i in range(1, Nt) if i % n_maps == 0]
    ns.append(Nt)
    jet_vals = np.zeros([len(ns), 3, 4, len(mesh.x)])
    jet_vals[0] = ev.identity_jets(mesh.vertices)
    map0 = ev.Map(mesh, jet_vals[0].copy())
    curr_map = map0
    submaps = Composite(mesh, jet_vals, ns)

    tspan = np.linspace(0, T, Nt, endpoint=False)
    dt = tspan[1] - tspan[0]
    mesh_u = meshes.StructuredTriangulation(L)
    Tu = len(mesh_u.simplices)
    U = sp.VelocitySpline(mesh_u, np.zeros([3, 3, 4, len(mesh_u.vertices)]),
                          np.zeros([3, 3, 19, Tu]), ts=tspan[0:3])
    D_inv, L_plus, L_minus, L_z = sht.operator_tables(L)
    Phi, The = np.meshgrid(mesh_u.phi, mesh_u.theta)
    sample_pts = mesh_u.sample_points()

    def omega(X):
        angs = g.cart2sphere(X)
        angs = [angs[0].reshape([L + 1, 2 * L]), angs[1].reshape([L + 1, 2 * L])]
        return (vorticity(angs[0], angs[1]) + 2 * rot_rate * np.cos(angs[1])
                - 2 * rot_rate * np.cos(The))

    psi0 = vorticity(Phi, The)
    psi_lms = sht.inv_Laplacian(psi0, L, D_inv)
    temp = 0 * psi_lms[:]
    sht.project_stream(psi_lms, L, L_plus, L_minus, L_z, U, 0, temp)
    log("bootstrap: initial velocity")

    t = 0
    int0 = ev.advect_project_sphere(map0, ev.Euler_step_proj, t, dt, U, identity=True)
    psi_lms1 = sht.inv_Laplacian(omega(int0(sample_pts)), L, D_inv)
    sht.project_stream(psi_lms1, L, L_plus, L_minus, L_z, U, 1, temp)
    log("bootstrap: Euler step")
    U.nVs = 2
    int1 = ev.advect_project_sphere(map0, ev.improved_Euler_proj, t, dt, U, identity=True)
    psi_lms2 = sht.inv_Laplacian(omega(int1(sample_pts)), L, D_inv)
    sht.project_stream(psi_lms2, L, L_plus, L_minus, L_z, U, 1, temp)
    log("bootstrap: improved Euler step")
    int2 = ev.advect_project_sphere(int1, ev.RK4_proj, t + dt, dt, U)
    psi_lms3 = sht.inv_Laplacian(omega(int2(sample_pts)), L, D_inv)
    U.nVs = 3
    sht.project_stream(psi_lms3, L, L_plus, L_minus, L_z, U, 2, temp)
    log("bootstrap: RK4 step")
    curr_map = int2

    U.stepping = True
    identity = False
    count = 2
    for t in tspan[2:]:
        curr_map = ev.advect_project_sphere(curr_map, ev.RK4_proj, t, dt, U, identity=identity)
        XXn = submaps(curr_map(sample_pts).T).T
        psi_lms_n = sht.inv_Laplacian(omega(XXn), L, D_inv)
        sht.project_stream(psi_lms_n, L, L_plus, L_minus, L_z, U, 2, temp)
        U.ts = U.ts + dt
        count += 1
        identity = False
        if step_hook is not None:
            step_hook(count, curr_map, U)
        if count in ns:
            idx = ns.index(count)
            submaps.Nc = idx + 1
            submaps.assemble_coefficients(i=idx, vals=curr_map.vals[:])
            curr_map = map0
            identity = True
    return submaps
