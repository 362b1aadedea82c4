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
    # two evaluations (dual-basis vs Cramer barycentrics, FMA contraction vs
    # NumPy's separate roundings) reach the jets multiplied by 1/(4 eps) (the
    # 1e-9 jet tolerance) and the map values multiplied by the mesh spacing
    # h ~ 0.04: measured after 10 steps, max 6.2e-11 and median 2.8e-12 in
    # both frames (99.9th percentile 4e-11).  With the poles on the flow's
    # fixed axis the pole stencils are location ties (the flow fixes the pole
    # vertices); the measured differences near the poles are no larger.
    tol = 1e-10
    away = np.abs(q[sub][:, 2]) < np.cos(0.5)
    print(f"C1 {frame}: max |gpu - oracle| {err.max():.3e}, outside the polar caps {err[away].max():.3e}, "
          f"median {np.median(err):.3e}, 99.9th percentile {np.quantile(err, 0.999):.3e}", flush=True)
    assert err.max() < tol
    assert np.median(err) < 1e-11


def test_zoom_vorticity_matches_oracle():
    """zoom_vorticity (figures/zoom_sampling_figure.py:63-90) against the
    oracle: the same windows through the oracle's composition of the run's
    submaps (scipy find_simplex + NumPy PS evaluation) and the analytic
    vorticity, with the reference's theta centre (x0) and with y0."""
    from oracle import meshes as om, splines as osp, fields as ofl
    from cmm_s2.core import dynamical_fields as df, mesh_functions as mf
    from cmm_s2.core.spherical_simulations import BarotropicCMM
    ico = mf.icosahedral_mesh(3)
    mesh = mf.spherical_triangulation(ico.points, simplices=ico.simplices)
    sim = BarotropicCMM(24, 12, 12 / 24, mesh, df.rossby_wave, 2 * np.pi, n_maps=4)
    sim.run()
    assert sim.identity  # the run ends with a remap: the chain is the submaps
    oms = om.SphericalTriangulation(mesh.vertices)
    vals = sim.submaps.vals[:sim.submaps.Nc].cpu().numpy()
    levels = [osp.assemble(vals[i], oms.simplices, oms.ps_split) for i in range(sim.submaps.Nc)]
    ks, n, x0, y0 = (-2, -5), 20, 3.22055, 1.1963
    for same in (True, False):
        got = sim.zoom_vorticity(x0, y0, ks, n, same_center=same).cpu().numpy()
        for j, k in enumerate(ks):
            h = 2.0 ** k
            xs = np.linspace(x0 - h, x0 + h, n)
            ys = np.linspace((x0 if same else y0) - h, (x0 if same else y0) + h, n)
            PH, TH = np.meshgrid(xs, ys)
            q = np.stack([np.sin(TH) * np.cos(PH), np.sin(TH) * np.sin(PH), np.cos(TH)], -1).reshape(-1, 3)
            X = osp.compose_eval(oms.query, levels, q)
            ph = np.mod(np.arctan2(X[:, 1], X[:, 0]) + 2 * np.pi, 2 * np.pi)
            th = np.arctan2(np.hypot(X[:, 0], X[:, 1]), X[:, 2])
            ref = (ofl.rossby_wave(ph, th) + 2 * 2 * np.pi * np.cos(th)).reshape(n, n)
            assert np.abs(got[j] - ref).max() <= 1e-11 * np.abs(ref).max(), (same, k)
