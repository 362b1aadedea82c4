"""GPU parity: map Jacobians (eval_grad), scalar spline gradients and the
Jacobian of a composed chain against the oracle (restated
spherical_spline.py:73-111, 208-265, 700-720; pinned by finite differences
and exact quadratics in test_oracle_splines_evolution.py)."""
import numpy as np
import pytest
import torch

from helpers import perturbed_identity_jets, random_unit, rel_err

pytestmark = pytest.mark.gpu


def _pair(k):
    from cmm_s2.core import mesh_functions as mf
    from oracle import meshes as om
    ico = mf.icosahedral_mesh(k)
    pm = mf.spherical_triangulation(ico.points, simplices=ico.simplices)
    oms = om.SphericalTriangulation(ico.points)
    return pm, oms


def test_map_jacobian_matches_oracle():
    from cmm_s2.core.spherical_spline import sphere_diffeomorphism
    from oracle import splines as osp
    pm, oms = _pair(3)
    jets = perturbed_identity_jets(pm.vertices, 21, amp=0.05)
    g = sphere_diffeomorphism(pm, torch.from_numpy(jets).cuda())
    c = osp.assemble(jets, oms.simplices, oms.ps_split)
    q = random_unit(20000, 22)
    got = g.eval_grad(q)
    assert got.shape == (3, 3, len(q))
    assert rel_err(got, osp.map_grad(oms.query, c, q)) < 1e-10


@pytest.mark.parametrize("mesh_kind", ["ico", "mwss"])
def test_scalar_gradient_matches_oracle(mesh_kind):
    from cmm_s2.core import mesh_functions as mf
    from cmm_s2.core.spherical_spline import spherical_spline
    from oracle import meshes as om, splines as osp
    if mesh_kind == "ico":
        pm, om_ = _pair(3)
    else:
        pm, om_ = mf.structure_spherical_triangulation(16), om.StructuredTriangulation(16)
    V = np.asarray(pm.vertices)
    rng = np.random.default_rng(5)
    vals = rng.standard_normal((4, len(V)))
    vals[1:] -= np.einsum("in,ni->n", vals[1:], V)[None] * V.T
    sp = spherical_spline(pm, torch.from_numpy(vals).cuda())
    c = osp.assemble(vals, om_.simplices, om_.ps_split)
    q = random_unit(20000, 23)
    got = sp.gradient(q)
    assert rel_err(got, np.array(osp.grad_eval(om_.query, c, q))) < 1e-10


def test_composed_jacobian_chain_rule():
    from cmm_s2.core.spherical_spline import composite_sphere_diffeomorphism
    from oracle import splines as osp
    pm, oms = _pair(3)
    vals = np.stack([perturbed_identity_jets(pm.vertices, 30 + i, amp=0.04) for i in range(3)])
    comp = composite_sphere_diffeomorphism(pm, torch.from_numpy(vals).cuda(), None, [1, 2, 3])
    for i in range(3):
        comp.assemble_coefficients(i, torch.from_numpy(vals[i]).cuda())
    comp.Nc = 3
    q = random_unit(20000, 24)
    X, J = comp.jacobian(q)
    assert rel_err(X, comp(q)) < 1e-15
    # oracle: levels in application order 2, 1, 0
    x = q.copy()
    Jo = np.broadcast_to(np.eye(3)[:, :, None], (3, 3, len(q))) - np.einsum("nc,nd->cdn", q, q)
    for i in (2, 1, 0):
        c = osp.assemble(vals[i], oms.simplices, oms.ps_split)
        Jl = osp.map_grad(oms.query, c, x)
        Jo = np.einsum("ckn,kdn->cdn", Jl, Jo)
        x = osp.map_eval(oms.query, c, x).T.copy()
    assert rel_err(X, x) < 1e-12
    assert rel_err(J, Jo) < 1e-10
