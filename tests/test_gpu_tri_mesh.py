"""GPU parity: unstructured-mesh location, assembly, evaluation, composition
and the stencil/jet projection against the oracle (scipy find_simplex and the
NumPy restatement of spherical_spline.py)."""
import numpy as np
import pytest
import torch

from helpers import perturbed_identity_jets, random_unit, rel_err, tri_map

pytestmark = pytest.mark.gpu

T_POS = 1e-11   # positions / values, relative max-norm
T_JET = 1e-9    # jets (finite differences with eps = 1e-5)
T_TIE = 1e-12   # orientation-margin tie band of the triangle-index contract


@pytest.fixture(scope="module")
def meshes():
    from cmm_s2.core import mesh_functions as mf
    from oracle import meshes as om
    ico = mf.icosahedral_mesh(3)
    pm = mf.spherical_triangulation(ico.points, simplices=ico.simplices)
    oms = om.SphericalTriangulation(ico.points)
    return pm, oms, tri_map(pm.simplices, oms.simplices)


def test_mesh_matches_delaunay(meshes):
    pm, oms, mp = meshes
    assert len(np.unique(mp)) == len(pm.simplices) == len(oms.simplices)


def test_locate_bit_exact_outside_ties(meshes):
    from oracle import meshes as om
    pm, oms, mp = meshes
    q = np.concatenate([random_unit(50000, 1), pm.vertices, np.eye(3), -np.eye(3)])
    bcc_o, tri_o, _ = oms.query(q)
    assert (tri_o >= 0).all()
    tri_o = mp[tri_o]
    t, b = pm.locate(torch.from_numpy(q).cuda(), "N3")
    tri_g = t.cpu().numpy().astype(np.int64)
    assert (tri_g >= 0).all()
    margin = om.orientation_margin(pm.vertices, pm.simplices, tri_o, q)
    clear = margin > T_TIE
    assert (tri_g[clear] == tri_o[clear]).all(), "triangle index differs away from ties"
    # ties: the GPU triangle contains the point within the tie band
    m_g = om.orientation_margin(pm.vertices, pm.simplices, tri_g, q)
    assert (m_g > -T_TIE).all()
    # barycentrics reproduce q = sum b_i v_i
    vs = pm.vertices[pm.simplices[tri_g]]
    rec = np.einsum("ni,nij->nj", b.cpu().numpy(), vs)
    assert np.abs(rec - q).max() < 1e-14


def test_delaunay_row_order_is_the_references():
    """triangulation="delaunay" builds the mesh as the reference does
    (Delaunay of the origin and the points, mesh_functions.py:41-56): the
    simplices equal the oracle's row for row, and query()[1] returns the
    reference's triangle indices directly (no vertex-triple remap)."""
    from cmm_s2.core import mesh_functions as mf
    from oracle import meshes as om
    ico = mf.icosahedral_mesh(3)
    pm = mf.spherical_triangulation(ico.points, triangulation="delaunay")
    oms = om.SphericalTriangulation(ico.points)
    assert np.array_equal(pm.simplices, oms.simplices)
    assert pm.mesh_d is not None
    q = random_unit(20000, 17)
    bo, to, _ = oms.query(q)
    b, t, _ = pm.query(torch.from_numpy(q).cuda())
    t = t.cpu().numpy()
    clear = om.orientation_margin(pm.vertices, pm.simplices, to, q) > T_TIE
    assert (t[clear] == to[clear]).all()
    assert np.abs(b.cpu().numpy()[clear] - bo[clear]).max() < 1e-12


def test_assemble_and_eval(meshes):
    from oracle import splines as osp
    from cmm_s2.core.spherical_spline import sphere_diffeomorphism
    pm, oms, mp = meshes
    jets = perturbed_identity_jets(pm.vertices, 3)
    coeffs = torch.zeros(3, 19, pm.ntri, dtype=torch.float64, device="cuda")
    m = sphere_diffeomorphism(pm, torch.from_numpy(jets).cuda(), coeffs)
    # coefficients depend on each triangle's vertex order: compare on the
    # product's simplices; the evaluation below compares the spline itself
    # against the Delaunay-ordered oracle mesh
    from oracle import meshes as om
    ref_p = osp.assemble(jets, pm.simplices, om.ps_split(pm.vertices, pm.simplices))
    assert rel_err(coeffs.cpu().numpy(), ref_p) < T_POS
    ref_c = osp.assemble(jets, oms.simplices, oms.ps_split)
    q = random_unit(40000, 4)
    ref = osp.map_eval(oms.query, ref_c, q)
    got = m(torch.from_numpy(q).cuda()).cpu().numpy()
    assert rel_err(got, ref) < T_POS
    # un-normalised evaluation = the raw BB polynomial
    raw_ref = osp.located_eval(oms.query, ref_c, q)
    raw = m.eval(torch.from_numpy(q).cuda(), "N3", normalize=False).cpu().numpy().T
    assert rel_err(raw, raw_ref) < T_POS


def test_compose_chain(meshes):
    from oracle import splines as osp
    from cmm_s2.core.spherical_spline import composite_sphere_diffeomorphism
    pm, oms, mp = meshes
    depth = 4
    vals = np.stack([perturbed_identity_jets(pm.vertices, 10 + i) for i in range(depth)])
    comp = composite_sphere_diffeomorphism(pm, torch.from_numpy(vals).cuda(), None, list(range(depth)))
    for i in range(depth):
        comp.assemble_coefficients(i, torch.from_numpy(vals[i]).cuda())
    comp.Nc = depth
    levels = [osp.assemble(vals[i], oms.simplices, oms.ps_split) for i in range(depth)]
    q = random_unit(30000, 5)
    ref_all = osp.compose_eval(oms.query, levels, q, eval_all=True)
    got = comp(torch.from_numpy(q).cuda()).cpu().numpy()
    assert rel_err(got, ref_all[-1]) < T_POS
    outs = comp.eval_all(torch.from_numpy(q).cuda())
    for a, b in zip(outs, ref_all):
        assert rel_err(a.cpu().numpy(), b) < T_POS


def test_stencil_eval_and_projection(meshes):
    from oracle import splines as osp, evolution as oev, meshes as om
    from cmm_s2.core.spherical_spline import sphere_diffeomorphism
    from cmm_s2.core import evolution_functions as evol
    pm, oms, mp = meshes
    jets = perturbed_identity_jets(pm.vertices, 6)
    m = sphere_diffeomorphism(pm, torch.from_numpy(jets).cuda())
    ref_c = osp.assemble(jets, oms.simplices, oms.ps_split)
    # feet: vertices / stencil points rotated about a generic axis (a rotation
    # about z fixes the poles ON mesh vertices: a tie where any incident
    # triangle is a valid location and the stencil extrapolation differs at
    # O(eps^2) between them)
    from oracle.geometry import rotation_matrix
    R = rotation_matrix([0.3, -0.5, 0.8], 0.01)
    yv = (R @ pm.vertices.T)
    st = np.einsum("ij,jkn->ikn", R, pm.s_pts)
    ref = osp.stencil_eval(oms.query, ref_c, yv, st.copy())
    got = m.stencil_eval(torch.from_numpy(yv).cuda(), torch.from_numpy(st.copy()).cuda()).cpu().numpy()
    assert rel_err(got, ref) < T_POS
    # stencil_pts of the product equal the oracle's restatement
    s_o, tv_o = om.stencil_points(pm.vertices)
    assert np.abs(pm.s_pts - s_o).max() < 1e-15
    # projection to jets
    class _I:  # oracle interpolant stub carrying the mesh frames
        pass
    it = _I()
    it.vals = jets.copy()
    it.mesh = oms
    oms_tv = oms.tan_vects
    oms.tan_vects = tv_o
    ref_map = oev.spline_proj_sphere(it, ref.copy())
    oms.tan_vects = oms_tv
    got_map = evol.spline_proj_sphere(m, torch.from_numpy(ref.copy()).cuda())
    gv = got_map.vals.cpu().numpy()
    assert rel_err(gv[:, 0], ref_map.vals[:, 0]) < T_POS
    assert rel_err(gv[:, 1:], ref_map.vals[:, 1:]) < T_JET


def test_locate_large_mesh_against_kdtree():
    """ico-6 (81920 triangles), 200k points: contract checked against the
    oracle's exact containment search."""
    from cmm_s2.core import mesh_functions as mf
    from oracle import meshes as om
    ico = mf.icosahedral_mesh(6)
    pm = mf.spherical_triangulation(ico.points, simplices=ico.simplices)
    q = random_unit(200000, 7)
    t_o, m_o = om.locate_kdtree(pm.vertices, pm.simplices, q)
    t, _ = pm.locate(torch.from_numpy(q).cuda(), "N3")
    t = t.cpu().numpy().astype(np.int64)
    clear = om.orientation_margin(pm.vertices, pm.simplices, t_o, q) > T_TIE
    assert (t[clear] == t_o[clear]).all()
    assert (om.orientation_margin(pm.vertices, pm.simplices, t, q) > -T_TIE).all()


def test_tiled_grid_composition_is_identical():
    """cmm_tri_compose_grid walks a ring-major grid in tiles: same bits as the
    point-order composition for every tile shape, including ragged edges
    (ring count and length not multiples of the tile), the intermediates,
    NaN points, and waves whose points are spread over many triangles
    (ico-6 mesh, L=24 grid)."""
    from cmm_s2.core import mesh_functions as mf
    from cmm_s2.core import spherical_spline as ss
    from helpers import perturbed_identity_jets
    for ico_level, Ls in ((3, (24, 37, 96)), (6, (24,))):
        ico = mf.icosahedral_mesh(ico_level)
        pm = mf.spherical_triangulation(ico.points, simplices=ico.simplices)
        maps = [ss.sphere_diffeomorphism(pm, torch.from_numpy(perturbed_identity_jets(pm.vertices, 40 + i)).cuda())
                for i in range(3)]
        levels = [m.packed for m in maps]
        for L in Ls:
            grid = mf.structure_spherical_triangulation(L)
            q = grid.sample_points().clone()
            q[5] = float("nan")
            q[2 * L + 7] = float("nan")
            n = q.shape[0]
            inter_a = torch.empty(3, n, 3, dtype=torch.float64, device="cuda")
            a = ss.compose_packed(pm, levels, q, "N3", inter=inter_a)
            assert torch.isnan(a[5]).all() and torch.isfinite(a[6]).all()
            key = (pm._h, L + 1, 2 * L, len(levels), False)
            for rows in (1, 2, 4, 8, 16):
                ss._tiles[key] = rows
                inter_b = torch.full_like(inter_a, 7.0)
                b = ss.compose_packed(pm, levels, q, "N3", inter=inter_b, grid=(L + 1, 2 * L))
                assert torch.equal(a.nan_to_num(9.0), b.nan_to_num(9.0)), (ico_level, L, rows)
                assert torch.equal(inter_a.nan_to_num(9.0), inter_b.nan_to_num(9.0)), (ico_level, L, rows)
            ss._tiles.pop(key)
            # the first level at stored locations: the same bits
            blk, cb = ss.split_locate(pm, q)
            assert int(blk[5]) == -1 and int(blk[2 * L + 7]) == -1 and bool((blk[6:2 * L] >= 0).all())
            tri = pm.locate(q)[0]
            assert torch.equal((blk // 6).long(), torch.where(tri >= 0, tri, -1).long().where(blk >= 0, -1))
            keyl = key[:-1] + (True,)
            for rows in (1, 4, 16):
                ss._tiles[keyl] = rows
                inter_c = torch.full_like(inter_a, 7.0)
                c = ss.compose_packed(pm, levels, q, "N3", inter=inter_c, grid=(L + 1, 2 * L), located=(blk, cb))
                assert torch.equal(a.nan_to_num(9.0), c.nan_to_num(9.0)), (ico_level, L, rows, "located")
                assert torch.equal(inter_a.nan_to_num(9.0), inter_c.nan_to_num(9.0)), (ico_level, L, rows)
            ss._tiles.pop(keyl)
