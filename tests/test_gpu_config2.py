"""GPU parity at the full barotropic configuration: icosahedral map mesh
level 7 (163,842 vertices, 327,680 triangles), MWSS velocity bandlimit
L = 512, Omega = 2 pi, initial vorticity a seeded l^-3 Gaussian field
(1 <= l <= 64, seed 2) as a scalar spline on the L = 1024 MWSS mesh,
100 RK4 steps of dt = 1/512 with a remap every 20 (composition depth 5),
composed map on the L = 2048 MWSS grid (8,392,704 points).

Against the oracle: triangle indices of the ico-7 mesh (Qhull with "QJ",
the reference's find_simplex) on 200,000 random points and the whole L = 512
sample grid; the FP64 forward transform of the last step's vorticity grid
and the 12-field synthesis of its velocity operators; the depth-5 composed
map on every 16th point of the L = 2048 grid.  At full size: unit norm of
the composed map, its area factor (exactly 1 for the incompressible flow),
the drift of enstrophy and energy over the run."""
import numpy as np
import pytest
import torch

from helpers import random_unit, rel_err, tri_map

pytestmark = pytest.mark.gpu

T_POS = 1e-11
T_TIE = 1e-12


@pytest.fixture(scope="module")
def run():
    from cmm_s2.core import mesh_functions as mf, spherical_harmonic_tools as st, dynamical_fields as df
    from cmm_s2.core.spherical_simulations import BarotropicCMM
    from oracle import meshes as om
    print("config 2: ico-7 mesh, L=512, 100 steps", flush=True)
    ico = mf.icosahedral_mesh(7)
    mesh = mf.spherical_triangulation(ico.points, simplices=ico.simplices)
    coeffs = df.random_sh_coefficients(2, 64, 1024)
    w0 = st.project_onto_S12(df.lm_to_full(coeffs, 1024), 1024, mf.structure_spherical_triangulation(1024))
    sim = BarotropicCMM(512, 100, 100 / 512, mesh, w0, 2 * np.pi, 20)
    sim.run()
    torch.cuda.synchronize()
    oms = om.SphericalTriangulation(mesh.vertices, qhull_options="QJ")
    return sim, mesh, oms


def test_ico7_triangle_indices(run):
    """Bit-exact triangle index (oracle rows mapped by vertex triples) for
    every query whose orientation margin in the oracle's triangle exceeds
    1e-12; inside the tie band the GPU's triangle contains the point."""
    from oracle import meshes as om
    sim, mesh, oms = run
    mp = tri_map(mesh.simplices, oms.simplices)
    assert len(np.unique(mp)) == mesh.ntri == 327680
    grid = sim.samples.cpu().numpy()
    for q in (random_unit(200000, 71), grid):
        _, tri_o, _ = oms.query(q)
        assert (tri_o >= 0).all()
        tri_o = mp[tri_o]
        t, b = mesh.locate(torch.from_numpy(q).cuda(), "N3")
        tri_g = t.cpu().numpy().astype(np.int64)
        clear = om.orientation_margin(mesh.vertices, mesh.simplices, tri_o, q) > T_TIE
        assert clear.mean() > 0.99
        assert (tri_g[clear] == tri_o[clear]).all()
        assert (om.orientation_margin(mesh.vertices, mesh.simplices, tri_g, q) > -T_TIE).all()
        # barycentrics reproduce q = sum b_i v_i; the dual basis of a triangle
        # of edge h ~ 0.01 has norm ~ 1/h^2 in its ill-conditioned direction,
        # so rounding reaches ~1e-12
        vs = mesh.vertices[mesh.simplices[tri_g]]
        assert np.abs(np.einsum("ni,nij->nj", b.cpu().numpy(), vs) - q).max() < 1e-11


def _packed_of(plan, a):
    """[l, m] complex (nf, L, L) -> packed (nlm, nf, 2)"""
    v = a[:, plan.l_of, plan.m_of]                     # (nf, nlm)
    return np.stack([v.real.T, v.imag.T], axis=-1)


def test_L512_transforms_against_oracle(run):
    """The last step's vorticity grid (not band-limited: w0 o X sampled on the
    grid) through the FP64 forward transform, and the 12 velocity-operator
    fields of that step through the FP64 synthesis, against the oracle's
    transforms: 1e-11 relative."""
    from oracle import sht as osh
    sim, _, _ = run
    L = 512
    w = sim.w.reshape(L + 1, 2 * L).cpu().numpy()
    ref = _packed_of(sim.plan, osh.forward_lm(w, L)[None])
    assert rel_err(sim.w_lm.cpu().numpy(), ref) < T_POS
    ops = sim.ops.cpu().numpy()                            # (nlm, 12, 2)
    a = np.zeros((12, L, L), complex)
    a[:, sim.plan.l_of, sim.plan.m_of] = (ops[:, :, 0] + 1j * ops[:, :, 1]).T
    ref = osh.inverse_lm(a, L)
    assert rel_err(sim.fields.cpu().numpy(), ref) < T_POS


def test_depth5_composition_on_the_L2048_grid(run):
    """The composed map (5 submaps) on every 16th point of the L = 2048 MWSS
    grid (524,544 points, grid order) against the oracle's composition
    (Qhull find_simplex + NumPy PS evaluation) of the same submap jets."""
    from cmm_s2.core import mesh_functions as mf
    from oracle import splines as osp
    sim, mesh, oms = run
    assert sim.submaps.Nc == 5
    grid = mf.structure_spherical_triangulation(2048)
    q = grid.sample_points()[::16].contiguous()
    got = sim.submaps(q).cpu().numpy()
    vals = sim.submaps.vals[:5].cpu().numpy()
    levels = [osp.assemble(vals[i], oms.simplices, oms.ps_split) for i in range(5)]
    ref = osp.compose_eval(oms.query, levels, q.cpu().numpy())
    assert rel_err(got, ref) < T_POS


def test_full_size_properties(run):
    """Whole L = 2048 grid: the composed map lands on the unit sphere, its
    area factor (incompressible flow: exactly 1) deviates by the spline
    error only, and the absolute vorticity's enstrophy and energy on the
    L = 1024 grid drift by less than 1e-6 over the 100 steps."""
    from cmm_s2.core import mesh_functions as mf
    sim, mesh, _ = run
    grid = mf.structure_spherical_triangulation(2048)
    q = grid.sample_points()
    X, J = sim.submaps.jacobian(q)
    assert X.shape == (8392704, 3)
    assert float((X.norm(dim=1) - 1).abs().max()) < 1e-14
    a = torch.cross(q, torch.tensor([0.3, -0.5, 0.8], dtype=q.dtype, device=q.device).expand_as(q), dim=1)
    t1 = a / a.norm(dim=1, keepdim=True)
    t2 = torch.cross(q, t1, dim=1)
    u = (J * t1.t()[None]).sum(dim=1).t()
    v = (J * t2.t()[None]).sum(dim=1).t()
    f = torch.cross(u, v, dim=1).norm(dim=1)
    dev = (f - 1).abs()
    assert float(dev.max()) < 5e-3 and float((dev ** 2).mean().sqrt()) < 2e-4
    cons = sim.conservation_errors(1024)
    assert cons["enstrophy_rel_err"] < 1e-6 and cons["energy_rel_err"] < 1e-6, cons
