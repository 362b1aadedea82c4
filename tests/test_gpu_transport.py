"""GPU parity of the linear-transport path (tests/test_spherical_transport.py
of the reference): the analytic velocity fields and tracer initial
conditions of dynamical_fields.py, caller-provided velocity / vorticity
callables, compose_maps, and the transport loop itself against the oracle."""
import numpy as np
import pytest
import torch

from helpers import perturbed_identity_jets, random_unit, rel_err

pytestmark = pytest.mark.gpu


def _rotation(axis, angle):
    a = np.asarray(axis, float) / np.linalg.norm(axis)
    K = np.array([[0, -a[2], a[1]], [a[2], 0, -a[0]], [-a[1], a[0], 0]])
    return np.eye(3) + np.sin(angle) * K + (1 - np.cos(angle)) * K @ K


def _rotated_mesh(k):
    """ico-k mesh turned off the coordinate axes: the vortex and deformation
    flows fix the poles, and a mesh vertex on a fixed point makes the
    stencil evaluation there a location tie"""
    from cmm_s2.core import mesh_functions as mf
    from oracle import meshes as om
    ico = mf.icosahedral_mesh(k)
    pts = ico.points @ _rotation([1, 2, 3], 0.37).T
    pm = mf.spherical_triangulation(pts, simplices=ico.simplices)
    oms = om.SphericalTriangulation(pts)
    oms.s_pts, oms.tan_vects = om.stencil_points(pm.vertices)
    return pm, oms


@pytest.mark.parametrize("vel", ["u_deform", "u_static_vortex", "u_rotating_vortex", "equator_vortex", "u_div"])
def test_transport_velocities(vel):
    from cmm_s2.core import dynamical_fields as df, evolution_functions as ev
    from oracle import fields as ofl, evolution as oev
    x = random_unit(20000, 21).T.copy()
    V, Vo = getattr(df, vel), getattr(ofl, vel)
    for t in (0.0, 0.3, 0.77):
        got = V(t, 0.01, torch.from_numpy(x).cuda()).cpu().numpy()
        assert rel_err(got, Vo(t, 0.01, x)) < 1e-12, t
        got = ev.RK4_proj(t, 0.01, torch.from_numpy(x).cuda(), V).cpu().numpy()
        assert rel_err(got, oev.RK4_proj(t, 0.01, x, Vo)) < 1e-13, t


def test_tracer_initial_conditions():
    from cmm_s2.core import dynamical_fields as df
    from oracle import fields as ofl
    rng = np.random.default_rng(22)
    phi = rng.uniform(0, 2 * np.pi, 40000)
    th = np.arccos(rng.uniform(-1, 1, 40000))
    # concentrate points on the bells / cylinders
    phi[:20000] = np.pi + rng.choice([-1, 1], 20000) * np.pi / 6 + rng.uniform(-0.6, 0.6, 20000)
    th[:20000] = np.pi / 2 + rng.uniform(-0.6, 0.6, 20000)
    got = df.cosine_bell(phi, th)
    assert np.abs(got - ofl.cosine_bell(phi, th)).max() < 1e-13
    got = df.IC_SC(phi, th)
    ref = ofl.IC_SC(phi, th)
    assert set(np.unique(got).round(12)) == {0.1, 1.0}
    assert (got != ref).sum() == 0
    X = np.stack([np.sin(th) * np.cos(phi), np.sin(th) * np.sin(phi), np.cos(th)])
    for t in (0.0, 0.4):
        assert np.abs(df.static_vort_IC(t, X) - ofl.static_vort_IC(t, X)).max() < 1e-12
        assert np.abs(df.rotating_vort_IC(t, X) - ofl.rotating_vort_IC(t, X)).max() < 1e-12
    Xs = X[:, :4000]
    ref = ofl.random_initial_condition(Xs)
    assert np.abs(df.random_initial_condition(Xs) - ref).max() < 1e-12 * np.abs(ref).max()


def test_plugin_velocity_callables():
    """A caller's V(t, dt, X): a NumPy function (the oracle's own field,
    evaluated on the host after one warning) and a torch function (evaluated
    on the device) integrate to the device kinds' results; the map update
    takes them too."""
    from cmm_s2.core import dynamical_fields as df, evolution_functions as ev, utils
    from cmm_s2.core.spherical_simulations import identity_jets
    from cmm_s2.core.spherical_spline import sphere_diffeomorphism
    from oracle import fields as ofl
    x = torch.from_numpy(random_unit(5000, 23).T.copy()).cuda()
    host_v = lambda t, dt, X: ofl.u_static_vortex(t, dt, X)
    with pytest.warns(RuntimeWarning):
        a = ev.RK4_proj(0.2, 0.01, x, host_v)
    assert rel_err(a.cpu().numpy(), ev.RK4_proj(0.2, 0.01, x, df.u_static_vortex).cpu().numpy()) < 1e-13
    torch_v = lambda t, dt, X: torch.stack(utils.cross(X, [0.0, 0.0, -2 * np.pi]))
    b = ev.improved_Euler_proj(0.1, 0.02, x, torch_v)
    assert rel_err(b.cpu().numpy(), ev.improved_Euler_proj(0.1, 0.02, x, df.u_sbr).cpu().numpy()) < 1e-14
    pm, _ = _rotated_mesh(3)
    g1 = g2 = sphere_diffeomorphism(pm, identity_jets(pm))
    g1 = ev.advect_project_sphere(g1, ev.RK4_proj, 0.0, 0.05, host_v, identity=True)
    j1 = g1.vals.clone()
    g2 = ev.advect_project_sphere(sphere_diffeomorphism(pm, identity_jets(pm)), ev.RK4_proj, 0.0, 0.05,
                                  df.u_static_vortex, identity=True)
    assert rel_err(j1[:, 0].cpu().numpy(), g2.vals[:, 0].cpu().numpy()) < 1e-13


def test_plugin_vorticity_callable_in_the_driver():
    """BarotropicCMM with the reference-style NumPy vorticity function (the
    oracle's rossby_wave) as the initial field: the run equals the run with
    the device kind to rounding."""
    from cmm_s2.core import mesh_functions as mf, dynamical_fields as df
    from cmm_s2.core.spherical_simulations import BarotropicCMM
    from oracle import fields as ofl
    ico = mf.icosahedral_mesh(2)
    pm = mf.spherical_triangulation(ico.points, simplices=ico.simplices)
    runs = []
    for w in (df.rossby_wave, lambda phi, theta: ofl.rossby_wave(phi, theta)):
        sim = BarotropicCMM(16, 12, 0.12, pm, w, 2 * np.pi, n_maps=5)
        with pytest.warns(RuntimeWarning) if w is not df.rossby_wave else _nullcontext():
            sim.run()
        runs.append(sim)
    q = torch.from_numpy(random_unit(5000, 24)).cuda()
    a, b = runs[0].submaps(q).cpu().numpy(), runs[1].submaps(q).cpu().numpy()
    assert rel_err(a, b) < 1e-11


class _nullcontext:
    def __enter__(self):
        return self

    def __exit__(self, *a):
        return False


def test_compose_maps_order():
    """compose_maps(remaps, evals, current) applies (remaps + current)[::-1]
    (evolution_functions.py:147-153): the current map first, then the
    remaps last to first, against the oracle's composition of the same
    maps."""
    from cmm_s2.core import evolution_functions as ev, mesh_functions as mf
    from cmm_s2.core.spherical_spline import sphere_diffeomorphism
    from oracle import evolution as oev, meshes as om
    ico = mf.icosahedral_mesh(3)
    pm = mf.spherical_triangulation(ico.points, simplices=ico.simplices)
    oms = om.SphericalTriangulation(ico.points)
    jets = [perturbed_identity_jets(pm.vertices, s, amp=0.05) for s in (31, 32, 33)]
    maps = [sphere_diffeomorphism(pm, torch.from_numpy(j).cuda()) for j in jets]
    omaps = [oev.Map(oms, j) for j in jets]
    q = random_unit(20000, 25)
    got = ev.compose_maps(maps[:2], q, [maps[2]])
    ref = oev.compose_maps(omaps[:2], q, [omaps[2]])
    assert got.shape == (len(q), 3)
    assert rel_err(got, ref) < 1e-12
    manual = omaps[0](omaps[1](omaps[2](q).T).T).T
    assert rel_err(got, manual) < 1e-12
    # without the current map, and a different order gives a different map
    got2 = ev.compose_maps(maps[:2], q)
    assert rel_err(got2, oev.compose_maps(omaps[:2], q)) < 1e-12
    assert rel_err(ev.compose_maps(maps[1::-1], q), got2) > 1e-4


def _test_func(X):
    x, y, z = X[:, 0], X[:, 1], X[:, 2]
    return np.sin(x ** 2) * (10 * np.cos(y - 2)) + np.cos(10 * z ** 5) + 2 * np.sin(x * 3 * (z - 0.1) + y ** 7)


def _eval_pts(n):
    phi = np.linspace(0, 2 * np.pi, n, endpoint=False)
    th = np.linspace(0, np.pi, n, endpoint=True)
    P, T = np.meshgrid(phi, th)
    return np.stack([np.sin(T) * np.cos(P), np.sin(T) * np.sin(P), np.cos(T)], -1).reshape(-1, 3)


def _transport(pm, V, T, Nt):
    """tests/test_spherical_transport.py:60-117 (no remapping) on the device"""
    from cmm_s2.core import evolution_functions as ev
    from cmm_s2.core.spherical_simulations import identity_jets
    from cmm_s2.core.spherical_spline import sphere_diffeomorphism
    mapping = sphere_diffeomorphism(pm, identity_jets(pm))
    tspan = np.linspace(0, T, Nt, endpoint=False)
    dt = tspan[1] - tspan[0]
    identity = True
    for t in tspan:
        mapping = ev.advect_project_sphere(mapping, ev.RK4_proj, t, dt, V, identity)
        identity = False
    return mapping


def test_transport_loop_matches_oracle():
    """The reference's transport loop with u_deform_rot (T = 1, Nt = 2^j + 10
    RK4 steps, no remapping) on ico-2 and ico-3: the final map on the
    reference's (phi, theta) evaluation grid against the oracle's loop, and
    the tracer error max |f(X_T(x)) - f(x)| of its test function (the flow
    returns to the identity at T = 1) falling with refinement."""
    from cmm_s2.core import dynamical_fields as df
    from oracle import evolution as oev, fields as ofl
    q = _eval_pts(60)
    errs = []
    for j in (2, 3):
        pm, oms = _rotated_mesh(j)
        Nt = 2 ** j + 10
        g = _transport(pm, df.u_deform_rot, 1.0, Nt)
        o = oev.transport(oms, ofl.u_deform_rot, 1.0, Nt)
        got = g(torch.from_numpy(q).cuda()).cpu().numpy()
        ref = o(q)
        # jets are eps = 1e-5 differences of stencil values: rounding times
        # 1/(4 eps) per step, carried into values by the mesh spacing
        assert rel_err(got, ref) < 1e-9, j
        e_gpu = np.abs(_test_func(got.T) - _test_func(q)).max()
        e_ref = np.abs(_test_func(ref.T) - _test_func(q)).max()
        assert abs(e_gpu - e_ref) < 1e-8 * max(1.0, e_ref)
        errs.append(e_gpu)
    for j in (4, 5):
        pm, _ = _rotated_mesh(j)
        g = _transport(pm, df.u_deform_rot, 1.0, 2 ** j + 10)
        X = g(torch.from_numpy(q).cuda()).cpu().numpy()
        errs.append(np.abs(_test_func(X.T) - _test_func(q)).max())
    assert all(b < a for a, b in zip(errs, errs[1:])), errs
    assert errs[-1] < 0.05 * errs[0], errs


def test_static_vortex_tracer_converges():
    """u_static_vortex carries static_vort_IC exactly: the tracer of the
    transported map, static_vort_IC(0, X_T(x)), approaches
    static_vort_IC(T, x) as the mesh is refined (device fields, device
    tracer sampling of the map's positions)."""
    from cmm_s2.core import dynamical_fields as df
    q = _eval_pts(80)
    exact = df.static_vort_IC(1.0, q.T.copy())
    errs = []
    for j in (3, 4, 5):
        pm, _ = _rotated_mesh(j)
        g = _transport(pm, df.u_static_vortex, 1.0, 2 ** j + 10)
        X = g(torch.from_numpy(q).cuda())                     # (3, N) device
        c = df.static_vort_IC.at(0.0).sample_xyz(X.t().contiguous()).cpu().numpy()
        errs.append(np.abs(c - exact).max())
    # measured 0.23, 0.067, 0.0105 at ico-3, 4, 5 (the vortex's sharp core)
    assert errs[1] < 0.5 * errs[0] and errs[2] < 0.5 * errs[1], errs
    assert errs[-1] < 0.1 * errs[0], errs
