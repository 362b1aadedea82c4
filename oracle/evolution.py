"""Characteristic time stepping and map algebra (oracle).
Restates cmm_s2/core/evolution_functions.py.

Pinned by tests/test_oracle_evolution.py: solid-body rotation known answers
(the projected RK4 step converges to the exact rotation at order 4 in dt, the
Euler / improved Euler steps at orders 1 / 2); spline_proj_sphere of exact
stencil feet of a rotation reproduces the rotation's jets (value R v, J^T =
(R (I - v v^T))^T) up to the O(eps^2) finite-difference error; compose of
two rotation maps equals the product rotation.
"""
import numpy as np

from . import geometry as g
from . import splines as sp


def div_norm(x):
    """evolution_functions.py:82-85."""
    normalize = 1 / np.sqrt(x[0] ** 2 + x[1] ** 2 + x[2] ** 2)
    return normalize[None, :] * x


def Euler_step_proj(t, dt, xn, V):
    """evolution_functions.py:19-23."""
    v1 = V(t, dt, xn)
    return div_norm(xn - dt * v1)


def improved_Euler_proj(t, dt, xn, V):
    """evolution_functions.py:25-33."""
    v1 = V(t, dt, xn)
    step1 = div_norm(xn - dt * v1)
    v2 = V(t + dt, dt, step1)
    return div_norm(xn - (dt / 2) * (v1 + v2))


def RK4_proj(t, dt, xn, V):
    """evolution_functions.py:50-67."""
    tpdt = t + dt
    v1 = V(tpdt, dt, xn)
    step1 = div_norm(xn - dt * v1 / 2)
    v2 = V(t + dt / 2, dt, step1)
    step2 = div_norm(xn - (dt / 2) * v2)
    v3 = V(t + 0.5 * dt, dt, step2)
    step3 = div_norm(xn - dt * v3)
    v4 = V(t, dt, step3)
    xn1 = xn - (dt / 6) * (v1 + 2 * v2 + 2 * v3 + v4)
    return div_norm(xn1)


def RK2_proj(t, dt, xn, V):
    """evolution_functions.py:70-80."""
    v1 = V(t + dt, dt, xn)
    step1 = div_norm(xn - (dt / 2) * v1)
    v2 = V(t + dt / 2, dt, step1)
    return div_norm(xn - dt * v2)


class Map:
    """Minimal sphere_diffeomorphism (spherical_spline.py:155-337): jets +
    coefficients on an icosahedral mesh with a ``query`` method."""

    def __init__(self, mesh, vals, coeffs=None):
        self.mesh = mesh
        self.vals = vals
        self.coeffs = sp.assemble(vals, mesh.simplices, mesh.ps_split)

    def __call__(self, q_pts):
        return sp.map_eval(self.mesh.query, self.coeffs, q_pts)

    def stencil_eval(self, q_pts, st_pts):
        return sp.stencil_eval(self.mesh.query, self.coeffs, q_pts, st_pts)


def identity_jets(vertices):
    """Identity map jets (spherical_simulations.py:165-169)."""
    jv = np.zeros([3, 4, len(vertices)])
    jv[:, 0, :] = vertices.T
    jv[:, 1:, :] = np.array([g.grad_x(vertices.T), g.grad_y(vertices.T), g.grad_z(vertices.T)])
    return jv


def spline_proj_sphere(interpolant, s_pts, eps=1e-5):
    """evolution_functions.py:90-113 -- stencil values -> jets -> new map."""
    jet_vals = interpolant.vals.copy()
    vals = (s_pts[:, 0, :] + s_pts[:, 1, :] + s_pts[:, 2, :] + s_pts[:, 3, :]) / 4
    vals = g.div_norm(vals.T)
    jet_vals[:, 0, :] = vals
    df_dx1 = (s_pts[:, 1, :] - s_pts[:, 0, :] + s_pts[:, 3, :] - s_pts[:, 2, :]) / (4 * eps)
    df_dx2 = (s_pts[:, 2, :] - s_pts[:, 0, :] + s_pts[:, 3, :] - s_pts[:, 1, :]) / (4 * eps)
    gammas = interpolant.mesh.tan_vects
    jet_vals[:, 1, :] = (df_dx1[0][:, None] * gammas[0] + df_dx2[0][:, None] * gammas[1]).T
    jet_vals[:, 2, :] = (df_dx1[1][:, None] * gammas[0] + df_dx2[1][:, None] * gammas[1]).T
    jet_vals[:, 3, :] = (df_dx1[2][:, None] * gammas[0] + df_dx2[2][:, None] * gammas[1]).T
    return Map(interpolant.mesh, jet_vals)


def advect(interp, integrator, t, dt, V, identity=False):
    """evolution_functions.py:117-133."""
    spts = np.array(interp.mesh.s_pts)
    verts0 = interp.mesh.vertices.T
    yn = integrator(t, dt, verts0, V)
    s = np.shape(spts)
    spts_n = integrator(t, dt, spts.reshape([3, s[1] * s[2]]), V)
    if not identity:
        return interp.stencil_eval(q_pts=yn, st_pts=spts_n.reshape(s))
    return spts_n.reshape(s)


def advect_project_sphere(interp, integrator, t, dt, V, identity=False):
    """evolution_functions.py:136-142."""
    s_evals = advect(interp, integrator, t, dt, V, identity)
    return spline_proj_sphere(interp, s_evals)


def compose_maps(remaps, evals, current=[]):
    """evolution_functions.py:147-153 -- evals is (N,3)... applied as maps(evals)."""
    rev_maps = (remaps + current)[::-1]
    evals_c = evals.copy()
    for i in range(len(rev_maps)):
        evals_c[:] = rev_maps[i](evals_c).T
    return evals_c


def transport(mesh, V, T, Nt, integrator=None):
    """The linear-transport loop of tests/test_spherical_transport.py:60-117
    without remapping (its default): identity jets, then Nt steps of
    advect_project_sphere over tspan = linspace(0, T, Nt, endpoint=False).
    Returns the final map."""
    integrator = RK4_proj if integrator is None else integrator
    mapping = Map(mesh, identity_jets(mesh.vertices))
    tspan = np.linspace(0, T, Nt, endpoint=False)
    dt = tspan[1] - tspan[0]
    identity = True
    for t in tspan:
        mapping = advect_project_sphere(mapping, integrator, t, dt, V, identity)
        identity = False
    return mapping
