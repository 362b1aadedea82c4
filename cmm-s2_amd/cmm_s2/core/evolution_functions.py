"""Characteristic time stepping and map algebra
(cmm_s2/core/evolution_functions.py API).

The integrators keep the reference signature ``integrator(t, dt, xn, V)``
(evolution_functions.py:19-80).  When V is a ``spline_interp_velocity`` or
one of the analytic fields of ``dynamical_fields`` the whole multi-stage step
(every stage's velocity location + time-interpolated evaluation +
projection) is ONE kernel launch per point set.  Any other callable
V(t, dt, X (3,N)) -> (3,N) (a caller's own field) is integrated stage by
stage on the device with the same formulas, V called on device tensors or,
when it only takes NumPy, on host copies.
"""
import numpy as np
import torch

from .. import _native as N
from . import _backend as B
from .spherical_spline import sphere_diffeomorphism, spline_interp_velocity, compose_packed

_STAGES = {
    N.INT_EULER: lambda t, dt: (t,),
    N.INT_IMPROVED_EULER: lambda t, dt: (t, t + dt),
    N.INT_RK4: lambda t, dt: (t + dt, t + dt / 2, t + 0.5 * dt, t),
    N.INT_RK2: lambda t, dt: (t + dt, t + dt / 2),
    N.INT_EVAL: lambda t, dt: (t,),
}


def velocity_descriptor(V, stage_times, dt):
    d = N.Velocity()
    if isinstance(V, spline_interp_velocity):
        d.kind = N.VEL_SPLINE
        nl = V.levels_used()
        d.nlev = nl
        d.mesh = V.mesh._h
        for k in range(nl):
            d.levels[k] = V.level_packed(k).data_ptr()
            if V.sparse:
                d.vslot[k] = V.level_slot(k)
                d.jets[k] = V.level_vals(k).data_ptr()
        if V.sparse:
            d.vbits = V._vbits.data_ptr()
            d.mark = V._mark.data_ptr()
            d.mark_every = int(V.mark_every)
        for s, ts in enumerate(stage_times):
            w = V.weights(ts, dt)
            for k in range(3):
                d.stage_w[s][k] = w[k]
    elif hasattr(V, "_cmm_kind"):
        d.kind = V._cmm_kind
        for k, p in enumerate(V._cmm_params):
            d.params[k] = p
    else:
        raise N.CmmError("device integrators need a spline_interp_velocity or a "
                         "cmm_s2.core.dynamical_fields velocity")
    for s, ts in enumerate(stage_times):
        d.stage_t[s] = ts
    return d


def is_device_velocity(V):
    return isinstance(V, spline_interp_velocity) or hasattr(V, "_cmm_kind")


def _norm(y):
    """div_norm (evolution_functions.py:82-85) on a (3, M) device tensor."""
    return (1 / torch.sqrt(y[0] ** 2 + y[1] ** 2 + y[2] ** 2))[None, :] * y


def integrate_plugin(kind, t, dt, x, V):
    """evolution_functions.py:19-80 for a caller's velocity callable V(t, dt,
    X (3,M)): the stages of integrator ``kind`` on device tensors."""
    f = lambda tt, y: B.call_plugin(V, tt, dt, y).reshape(3, -1)
    if kind == N.INT_EVAL:
        return f(t, x)
    if kind == N.INT_EULER:
        return _norm(x - dt * f(t, x))
    if kind == N.INT_IMPROVED_EULER:
        v1 = f(t, x)
        v2 = f(t + dt, _norm(x - dt * v1))
        return _norm(x - (dt / 2) * (v1 + v2))
    if kind == N.INT_RK2:
        v1 = f(t + dt, x)
        v2 = f(t + dt / 2, _norm(x - (dt / 2) * v1))
        return _norm(x - dt * v2)
    v1 = f(t + dt, x)
    v2 = f(t + dt / 2, _norm(x - dt * v1 / 2))
    v3 = f(t + 0.5 * dt, _norm(x - (dt / 2) * v2))
    v4 = f(t, _norm(x - dt * v3))
    return _norm(x - (dt / 6) * (v1 + 2 * v2 + 2 * v3 + v4))


def integrate(kind, t, dt, x, V, out=None):
    """One backward step of integrator ``kind`` for points x (3,M) on device."""
    m = x.shape[1]
    if out is None:
        out = B.empty(3, m)
    if not is_device_velocity(V):
        out.copy_(integrate_plugin(kind, t, dt, x, V))
        return out
    desc = velocity_descriptor(V, _STAGES[kind](t, dt), dt)
    N.call("cmm_advect_points", desc, kind, float(dt), x.data_ptr(), m, out.data_ptr(), B.stream())
    return out


def _wrap(kind):
    def step(t, dt, xn, V):
        x = B.to_dev(xn)
        return B.like(integrate(kind, t, dt, x, V), xn)
    return step


Euler_step_proj = _wrap(N.INT_EULER)
Euler_step_proj.__doc__ = "evolution_functions.py:19-23."
improved_Euler_proj = _wrap(N.INT_IMPROVED_EULER)
improved_Euler_proj.__doc__ = "evolution_functions.py:25-33."
RK4_proj = _wrap(N.INT_RK4)
RK4_proj.__doc__ = "evolution_functions.py:50-67."
RK2_proj = _wrap(N.INT_RK2)
RK2_proj.__doc__ = "evolution_functions.py:70-80."

_KIND = {Euler_step_proj: N.INT_EULER, improved_Euler_proj: N.INT_IMPROVED_EULER,
         RK4_proj: N.INT_RK4, RK2_proj: N.INT_RK2}


def div_norm(x):
    """evolution_functions.py:82-85 (host / torch helper)."""
    n = 1 / (x[0] ** 2 + x[1] ** 2 + x[2] ** 2) ** 0.5
    return n[None, :] * x


class _Work:
    """Reusable device buffers of one mesh for the map update."""

    def __init__(self, mesh):
        nv = mesh.nv
        self.yv = B.empty(3, nv)
        self.ys = B.empty(3, 4 * nv)
        self.jets = [B.empty(3, 4, nv), B.empty(3, 4, nv)]
        self.k = 0

    def next_jets(self):
        self.k ^= 1
        return self.jets[self.k]


def _work(mesh):
    """Map-update buffers of ``mesh`` for the current stream (simulations on
    different streams share the mesh tables, never these buffers)."""
    ws = getattr(mesh, "_cmm_work", None)
    if ws is None:
        ws = {}
        mesh._cmm_work = ws
    key = B.stream()
    w = ws.get(key)
    if w is None:
        w = _Work(mesh)
        if not torch.cuda.is_current_stream_capturing():  # graph-pool buffers stay with their graph
            ws[key] = w
    return w


def _feet(interp, integrator, t, dt, V, identity, w):
    mesh = interp.mesh
    if integrator in _KIND and not is_device_velocity(V):
        kind = _KIND[integrator]
        if not identity:
            w.yv.copy_(integrate_plugin(kind, t, dt, mesh.d_verts3, V))
        w.ys.copy_(integrate_plugin(kind, t, dt, mesh.d_spts, V))
    elif integrator in _KIND:
        kind = _KIND[integrator]
        desc = velocity_descriptor(V, _STAGES[kind](t, dt), dt)
        xv, xs = mesh.d_verts3, mesh.d_spts
        N.call("cmm_advect_stencil", desc, kind, float(dt), None if identity else xv.data_ptr(), xv.stride(0),
               xs.data_ptr(), xs.stride(0), mesh.nv, w.yv.data_ptr(), w.yv.stride(0), w.ys.data_ptr(),
               w.ys.stride(0), B.stream())
    else:  # a caller-provided integrator callable
        if not identity:
            w.yv.copy_(B.to_dev(integrator(t, dt, mesh.d_verts3, V)))
        w.ys.copy_(B.to_dev(integrator(t, dt, mesh.d_spts, V)))


def advect(interp, integrator, t, dt, V, identity=False):
    """evolution_functions.py:117-133: stencil values (3,4,Nv) of the updated map."""
    mesh = interp.mesh
    w = _work(mesh)
    _feet(interp, integrator, t, dt, V, identity, w)
    if identity:
        return w.ys.reshape(3, 4, mesh.nv).clone()
    return interp.stencil_eval(w.yv, w.ys.reshape(3, 4, mesh.nv))


def spline_proj_sphere(interpolant, s_pts, eps=1e-5):
    """evolution_functions.py:90-113: stencil values -> jets -> new map that
    takes over the old map's coefficient buffers (as the reference does)."""
    mesh = interpolant.mesh
    st = B.to_dev(s_pts).reshape(3, 4 * mesh.nv)
    jets = B.empty(3, 4, mesh.nv)
    N.call("cmm_tri_stencil_jets", mesh._h, None, None, 0, st.data_ptr(), st.stride(0),
           mesh.d_tan.data_ptr(), float(eps), jets.data_ptr(), None, B.stream())
    return sphere_diffeomorphism(mesh, jets, coeffs=interpolant._coeffs, packed=interpolant.packed)


def advect_project_sphere(interp, integrator, t, dt, V, identity=False):
    """evolution_functions.py:136-142, fused: integrate the vertex and stencil
    feet, evaluate the current map on them and project to jets in one
    stencil kernel, then assemble into the current map's buffer."""
    mesh = interp.mesh
    w = _work(mesh)
    _feet(interp, integrator, t, dt, V, identity, w)
    jets = w.next_jets()
    cur = None if identity else interp.packed.data_ptr()
    N.call("cmm_tri_stencil_jets", mesh._h, cur, w.yv.data_ptr(), w.yv.stride(0), w.ys.data_ptr(),
           w.ys.stride(0), mesh.d_tan.data_ptr(), 1e-5, jets.data_ptr(), None, B.stream())
    return sphere_diffeomorphism(mesh, jets, coeffs=None, packed=interp.packed)


def compose_maps(remaps, evals, current=[]):
    """evolution_functions.py:147-153: apply maps last-to-first to points
    evals (N,3) -> (N,3) in one fused launch."""
    maps = (list(remaps) + list(current))[::-1]
    q = B.to_dev(evals)
    if not maps:
        return B.like(q.clone(), evals)
    mesh = maps[0].mesh
    return B.like(compose_packed(mesh, [m.packed for m in maps], q, "N3"), evals)
