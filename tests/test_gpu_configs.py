"""GPU parity at the plain-CMM configuration: ico-5 map mesh (10,242 vertices),
Rossby-Haurwitz wavenumber-4 velocity (omega = K = 7.848e-6, R = 4,
Omega = 2 pi), identity start, 10 backward RK4 steps of dt = 0.01 with the time
grid of the reference's transport driver (tests/test_spherical_transport.py:
89-106: tspan = linspace(0, T, Nt, endpoint=False), identity on the first
step), then the map evaluated at 1,000,000 seeded uniform random points."""
import numpy as np
import pytest
import torch

from helpers import random_unit, rel_err

pytestmark = pytest.mark.gpu


def _rotation(axis, angle):
    a = np.asarray(axis, float) / np.linalg.norm(axis)
    K = np.array([[0, -a[2], a[1]], [a[2], 0, -a[0]], [-a[1], a[0], 0]])
    return np.eye(3) + np.sin(angle) * K + (1 - np.cos(angle)) * K @ K


@pytest.mark.parametrize("frame", ["poles_on_axis", "rotated"])
def test_plain_cmm_rh4_ico5(frame):
    from cmm_s2.core import mesh_functions as mf, evolution_functions as ev, dynamical_fields as df
    from cmm_s2.core.spherical_simulations import identity_jets
    from cmm_s2.core.spherical_spline import sphere_diffeomorphism
    from oracle import meshes as om, evolution as oev, fields as ofl
    ico = mf.icosahedral_mesh(5)
    pts = ico.points if frame == "poles_on_axis" else ico.points @ _rotation([1, 2, 3], 0.37).T
    pm = mf.spherical_triangulation(pts, simplices=ico.simplices)
    oms = om.SphericalTriangulation(pts, qhull_options="QJ")
    key = lambda S: {tuple(sorted(t)) for t in np.asarray(S).tolist()}
    assert key(oms.simplices) == key(pm.simplices)
    oms.s_pts, oms.tan_vects = om.stencil_points(pm.vertices)
    g = sphere_diffeomorphism(pm, identity_jets(pm))
    o = oev.Map(oms, oev.identity_jets(pm.vertices))
    tspan = np.linspace(0, 0.1, 10, endpoint=False)
    dt = tspan[1] - tspan[0]
    ident = True
    for t in tspan:
        g = ev.advect_project_sphere(g, ev.RK4_proj, t, dt, df.u_rh4, identity=ident)
        o = oev.advect_project_sphere(o, oev.RK4_proj, t, dt, ofl.u_rh4, identity=ident)
        ident = False
    q = random_unit(1_000_000, 1)
    got = g(torch.from_numpy(q).cuda()).cpu().numpy()      # (3, N), as sphere_diffeomorphism.__call__
    assert got.shape == (3, len(q)) and np.isfinite(got).all()
    assert np.abs(np.linalg.norm(got, axis=0) - 1).max() < 1e-14
    # the oracle on every 10th point (scipy find_simplex + NumPy)
    sub = slice(0, None, 10)
    ref = o(q[sub])
    err = np.abs(got[:, sub] - ref).max(axis=0)
    # Tolerance: every step rebuilds the map from jets that are eps = 1e-5
    # finite differences of stencil values, so round-off differences of the
    # two evaluations (Cramer vs dual-basis barycentrics) reach the jets
    # multiplied by 1/(4 eps) (the 1e-9 jet tolerance) and the map values
    # multiplied by the mesh spacing h ~ 0.04: 1e-10 after 10 steps (typical
    # differences are ~1e-12).
    tol = 1e-10
    if frame == "rotated":
        # no mesh vertex on the flow's fixed axis: no location ties
        assert err.max() < tol
    else:
        # The flow fixes both poles exactly (u = 0 there) and the poles are
        # mesh vertices, so the stencil evaluation at a pole is a location
        # tie: the oracle (Qhull's walk) and the GPU (first candidate) may
        # take different sub-triangles, whose polynomials differ at
        # O(eps^2) at the stencil points.  Each step carries that difference
        # one vertex ring further (~0.04 rad at ico-5), so after 10 steps it
        # stays inside caps of ~0.4 rad; outside them the maps agree to the
        # tolerance above.
        away = np.abs(q[sub][:, 2]) < np.cos(0.5)
        assert err[away].max() < tol
        assert err.max() < 1e-7
