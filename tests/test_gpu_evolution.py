"""GPU parity: backward integrators (spline and analytic velocities), the fused
map update and a scaled-down end-to-end barotropic run against the oracle."""
import numpy as np
import pytest
import torch

from helpers import random_unit, rel_err, tri_map

pytestmark = pytest.mark.gpu


def _velocity_pair(L, seed):
    from cmm_s2.core import mesh_functions as mf
    from cmm_s2.core.spherical_spline import spline_interp_velocity
    from oracle import meshes as om, splines as osp
    pm = mf.structure_spherical_triangulation(L)
    o = om.StructuredTriangulation(L)
    rng = np.random.default_rng(seed)
    vals = 0.5 * rng.standard_normal((3, 3, 4, pm.nv))
    U = spline_interp_velocity(pm, torch.from_numpy(vals).cuda(), None, ts=[0.0, 0.05, 0.1])
    Uo = osp.VelocitySpline(o, vals.copy(), np.zeros((3, 3, 19, pm.ntri)), ts=np.array([0.0, 0.05, 0.1]))
    for i in range(3):
        U.init_coeffs(i)
        Uo.init_coeffs(i)
    return U, Uo


@pytest.mark.parametrize("name", ["Euler_step_proj", "improved_Euler_proj", "RK4_proj", "RK2_proj"])
def test_integrators_spline_velocity(name):
    from cmm_s2.core import evolution_functions as ev
    from oracle import evolution as oev
    U, Uo = _velocity_pair(16, 1)
    x = random_unit(30000, 2).T.copy()
    for nVs, t in [(1, 0.0), (2, 0.05), (3, 0.1)]:
        U.nVs = Uo.nVs = nVs
        got = getattr(ev, name)(t, 0.05, torch.from_numpy(x).cuda(), U).cpu().numpy()
        ref = getattr(oev, name)(t, 0.05, x, Uo)
        assert rel_err(got, ref) < 1e-11


@pytest.mark.parametrize("vel", ["u_sbr", "u_deform_rot", "u_rh4"])
def test_analytic_velocities(vel):
    from cmm_s2.core import dynamical_fields as df, evolution_functions as ev
    from oracle import fields as ofl, evolution as oev
    x = random_unit(20000, 3).T.copy()
    V, Vo = getattr(df, vel), getattr(ofl, vel)
    got = V(0.3, 0.01, torch.from_numpy(x).cuda()).cpu().numpy()
    assert rel_err(got, Vo(0.3, 0.01, x)) < 1e-12
    got = ev.RK4_proj(0.3, 0.01, torch.from_numpy(x).cuda(), V).cpu().numpy()
    assert rel_err(got, oev.RK4_proj(0.3, 0.01, x, Vo)) < 1e-12


def test_stencil_advection_matches_pointwise():
    """cmm_advect_stencil (vertex + stencil feet interleaved in one launch)
    is the same arithmetic per point as cmm_advect_points: bit-identical."""
    from cmm_s2 import _native as N
    from cmm_s2.core import mesh_functions as mf, evolution_functions as ev
    from cmm_s2.core import _backend as B
    U, _ = _velocity_pair(16, 4)
    U.nVs = 3
    ico = mf.icosahedral_mesh(2)
    mesh = mf.spherical_triangulation(ico.points, simplices=ico.simplices)
    xv, xs = mesh.d_verts3, mesh.d_spts
    for kind in (N.INT_RK4, N.INT_EULER):
        ref_v = ev.integrate(kind, 0.1, 0.05, xv, U)
        ref_s = ev.integrate(kind, 0.1, 0.05, xs, U)
        for with_verts in (True, False):
            yv = torch.full_like(xv, -7.0)
            ys = torch.empty_like(xs)
            desc = ev.velocity_descriptor(U, ev._STAGES[kind](0.1, 0.05), 0.05)
            N.call("cmm_advect_stencil", desc, kind, 0.05, xv.data_ptr() if with_verts else None, xv.stride(0),
                   xs.data_ptr(), xs.stride(0), mesh.nv, yv.data_ptr(), yv.stride(0), ys.data_ptr(), ys.stride(0),
                   B.stream())
            torch.cuda.synchronize()
            assert torch.equal(ys, ref_s)
            if with_verts:
                assert torch.equal(yv, ref_v)
            else:
                assert bool((yv == -7.0).all())


@pytest.mark.parametrize("ico_level,L", [(5, 64), (6, 128)])
def test_stencil_advection_sparse_velocity_matches_pointwise(ico_level, L):
    """The RK4 feet of a run's map on its sparse velocity levels (blocks
    missing where no foot fell in the previous step, rebuilt from the jets
    during the evaluation): the vertex and stencil feet of one launch give
    the bits of the per-point kernel (cmm_advect_points) and of the
    stencil-only launch."""
    from cmm_s2 import _native as N
    from cmm_s2.core import mesh_functions as mf, evolution_functions as ev, dynamical_fields as df
    from cmm_s2.core import _backend as B
    from cmm_s2.core.spherical_simulations import BarotropicCMM
    ico = mf.icosahedral_mesh(ico_level)
    mesh = mf.spherical_triangulation(ico.points, simplices=ico.simplices)
    sim = BarotropicCMM(L, 8, 8 / L, mesh, df.rossby_wave.at(t=0.2), 2 * np.pi, n_maps=4)
    sim.bootstrap()
    for _ in range(3):
        sim.step()
    torch.cuda.synchronize()
    U = sim.U
    assert U.sparse
    xv, xs = mesh.d_verts3, mesh.d_spts
    t, dt = float(sim.tspan[sim.count]), sim.dt
    ref_v = ev.integrate(N.INT_RK4, t, dt, xv, U)
    outs = {}
    for with_verts in (True, False):
        yv = torch.full_like(xv, -7.0)
        ys = torch.empty_like(xs)
        desc = ev.velocity_descriptor(U, ev._STAGES[N.INT_RK4](t, dt), dt)
        N.call("cmm_advect_stencil", desc, N.INT_RK4, dt, xv.data_ptr() if with_verts else None, xv.stride(0),
               xs.data_ptr(), xs.stride(0), mesh.nv, yv.data_ptr(), yv.stride(0), ys.data_ptr(), ys.stride(0),
               B.stream())
        torch.cuda.synchronize()
        outs[with_verts] = (yv, ys)
    assert torch.isfinite(outs[True][1]).all()
    assert torch.equal(outs[True][1], outs[False][1])
    assert torch.equal(outs[True][0], ref_v)


def test_fused_map_update_matches_oracle():
    """advect_project_sphere with identity=True then False, analytic and
    spline velocities, against evolution_functions.py restated."""
    from cmm_s2.core import mesh_functions as mf, evolution_functions as ev, dynamical_fields as df
    from cmm_s2.core.spherical_simulations import identity_jets
    from cmm_s2.core.spherical_spline import sphere_diffeomorphism
    from oracle import meshes as om, evolution as oev, fields as ofl
    ico = mf.icosahedral_mesh(3)
    pm = mf.spherical_triangulation(ico.points, simplices=ico.simplices)
    oms = om.SphericalTriangulation(ico.points)
    oms.s_pts, oms.tan_vects = om.stencil_points(pm.vertices)
    mp = tri_map(pm.simplices, oms.simplices)
    m0 = sphere_diffeomorphism(pm, identity_jets(pm))
    o0 = oev.Map(oms, oev.identity_jets(pm.vertices))
    assert rel_err(m0.vals.cpu().numpy(), o0.vals) < 1e-15
    # tilted axes (alpha = 0.3): with alpha = 0 the flow fixes the poles,
    # which are mesh vertices, and every later stencil evaluation there is a
    # location tie
    Vg = df.DeviceVelocity("u_deform_rot", df.N.VEL_DEFORM_ROT, [1.0, 0.3])
    Vo = lambda t, dt, X: ofl.u_deform_rot(t, dt, X, alpha=0.3)
    g, o = m0, o0
    for k, ident in enumerate([True, False, False]):
        g = ev.advect_project_sphere(g, ev.RK4_proj, 0.1 * k, 0.05, Vg, identity=ident)
        o = oev.advect_project_sphere(o, oev.RK4_proj, 0.1 * k, 0.05, Vo, identity=ident)
        gv = g.vals.cpu().numpy()
        # jets come from eps = 1e-5 finite differences (round-off x 1/(4 eps));
        # a map rebuilt from them carries those jet errors times the mesh
        # spacing h (~0.1 on ico-3) into its values at the next step
        assert rel_err(gv[:, 0], o.vals[:, 0]) < 1e-10
        assert rel_err(gv[:, 1:], o.vals[:, 1:]) < 1e-8
    q = random_unit(20000, 4)
    assert rel_err(g(torch.from_numpy(q).cuda()).cpu().numpy(), o(q)) < 1e-9


def test_end_to_end_small_run():
    """ico-2 map, L=16 velocity, 12 steps with remaps every 5: composed map and
    vorticity on the sample grid against the oracle driver."""
    from cmm_s2.core import mesh_functions as mf, dynamical_fields as df
    from cmm_s2.core.spherical_simulations import BarotropicCMM
    from oracle import meshes as om, driver as odr, fields as ofl
    ico = mf.icosahedral_mesh(2)
    pm = mf.spherical_triangulation(ico.points, simplices=ico.simplices)
    oms = om.SphericalTriangulation(ico.points)
    oms.s_pts, oms.tan_vects = om.stencil_points(pm.vertices)
    L, Nt, T = 16, 12, 0.12
    sim = BarotropicCMM(L, Nt, T, pm, df.rossby_wave, 2 * np.pi, n_maps=5)
    sim.run()
    ref = odr.euler_simulation_sphere_remapping(L, Nt, T, oms, ofl.rossby_wave, 2 * np.pi, n_maps=5)
    assert sim.submaps.Nc == ref.Nc == 3
    q = random_unit(5000, 9)
    got = sim.submaps(torch.from_numpy(q).cuda()).cpu().numpy()
    assert rel_err(got, ref(q)) < 1e-9
