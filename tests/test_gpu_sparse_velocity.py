"""GPU: velocity levels assembled on a subset of triangles
(cmm_mwss_assemble_sparse, cmm_velocity.valid / jets / mark).  A missing
block is built from the level's jets with the assembly's arithmetic, so every
result must carry the bits of complete levels, whatever the assembled set."""
import numpy as np
import pytest
import torch

from helpers import random_unit

pytestmark = pytest.mark.gpu


def _velocity(L, seed):
    from cmm_s2.core import mesh_functions as mf
    from cmm_s2.core.spherical_spline import spline_interp_velocity
    pm = mf.structure_spherical_triangulation(L)
    rng = np.random.default_rng(seed)
    vals = 0.5 * rng.standard_normal((3, 3, 4, pm.nv))
    U = spline_interp_velocity(pm, torch.from_numpy(vals).cuda(), None, ts=[0.0, 0.05, 0.1])
    for i in range(3):
        U.init_coeffs(i)
    U.nVs = 3
    return U


def _sparse_copy(U, keep, seed):
    """U's levels with only a random fraction ``keep`` of each level's blocks
    valid; the others overwritten with NaN (they must never be read)."""
    from cmm_s2.core.spherical_spline import spline_interp_velocity
    V = spline_interp_velocity(U.mesh, U._vals.clone(), None, ts=U.ts)
    V._packed.copy_(U._packed)
    V._order = list(U._order)
    V.nVs = U.nVs
    V.sparse = True
    g = torch.Generator(device="cuda").manual_seed(seed)
    V._vbits.zero_()
    for k in range(3):
        v = (torch.rand(U.mesh.ntri, generator=g, device="cuda") < keep).to(torch.uint8)
        V._vbits |= v << k
        V._packed[k][v == 0] = float("nan")
    V._mark.zero_()
    return V


def _polar(mesh):
    """Triangles of the 4 rows next to either pole (always assembled)."""
    T, Nn = mesh.ntri, 2 * mesh.L
    t = torch.arange(T, device="cuda")
    return (t < Nn + 2 * Nn * 3) | (t >= T - Nn - 2 * Nn * 3)


@pytest.mark.parametrize("keep", [0.0, 0.3, 0.9])
@pytest.mark.parametrize("name", ["Euler_step_proj", "improved_Euler_proj", "RK4_proj"])
def test_missing_blocks_give_the_bits_of_complete_levels(name, keep):
    from cmm_s2.core import evolution_functions as ev
    U = _velocity(24, 1)
    V = _sparse_copy(U, keep, 7)
    x = torch.from_numpy(random_unit(40000, 2).T.copy()).cuda()
    for nVs, t in [(1, 0.0), (2, 0.05), (3, 0.1)]:
        U.nVs = V.nVs = nVs
        ref = getattr(ev, name)(t, 0.05, x, U)
        got = getattr(ev, name)(t, 0.05, x, V)
        torch.cuda.synchronize()
        assert torch.equal(got, ref), (nVs, t)
    # every evaluation marked its triangle: the feet's triangles of stage 1
    # (the start points) are among the marks
    from cmm_s2 import _native as N
    from cmm_s2.core import _backend as B
    tri = torch.empty(x.shape[1], dtype=torch.int32, device="cuda")
    N.call("cmm_mwss_locate", U.mesh._h, x.data_ptr(), x.shape[1], x.stride(1), x.stride(0), tri.data_ptr(),
           None, B.stream())
    assert bool((V._mark[tri.long()] == 1).all())


def test_stencil_feet_with_sparse_levels():
    """The fused vertex + stencil feet kernel (the loop's map update)."""
    from cmm_s2 import _native as N
    from cmm_s2.core import mesh_functions as mf, evolution_functions as ev
    from cmm_s2.core import _backend as B
    U = _velocity(32, 4)
    V = _sparse_copy(U, 0.25, 9)
    ico = mf.icosahedral_mesh(4)
    mesh = mf.spherical_triangulation(ico.points, simplices=ico.simplices)
    xv, xs = mesh.d_verts3, mesh.d_spts
    out = {}
    for name, W in (("full", U), ("sparse", V)):
        yv = torch.empty_like(xv)
        ys = torch.empty_like(xs)
        desc = ev.velocity_descriptor(W, ev._STAGES[N.INT_RK4](0.1, 0.05), 0.05)
        N.call("cmm_advect_stencil", desc, N.INT_RK4, 0.05, xv.data_ptr(), xv.stride(0), xs.data_ptr(),
               xs.stride(0), mesh.nv, yv.data_ptr(), yv.stride(0), ys.data_ptr(), ys.stride(0), B.stream())
        out[name] = (yv, ys)
    torch.cuda.synchronize()
    assert torch.equal(out["sparse"][0], out["full"][0])
    assert torch.equal(out["sparse"][1], out["full"][1])


def test_sparse_assembly_writes_exactly_the_marked_blocks():
    from cmm_s2 import _native as N
    from cmm_s2.core import _backend as B
    U = _velocity(24, 2)
    T = U.mesh.ntri
    full = U.level_packed(1).clone()
    packed = torch.full_like(full, float("nan"))
    g = torch.Generator(device="cuda").manual_seed(3)
    other = torch.randint(0, 256, (T,), generator=g, device="cuda", dtype=torch.int32).to(torch.uint8)
    slot = 5
    vbits = other.clone()
    mark = (torch.rand(T, generator=g, device="cuda") < 0.2).to(torch.uint8)
    sel = (mark == 1) | _polar(U.mesh)
    for bad in (-1, 8):
        with pytest.raises(N.CmmError):
            N.call("cmm_mwss_assemble_sparse", U.mesh._h, U.level_vals(1).data_ptr(), packed.data_ptr(),
                   vbits.data_ptr(), bad, -1, mark.data_ptr(), B.stream())
    N.call("cmm_mwss_assemble_sparse", U.mesh._h, U.level_vals(1).data_ptr(), packed.data_ptr(), vbits.data_ptr(),
           slot, -1, mark.data_ptr(), B.stream())
    torch.cuda.synchronize()
    assert torch.equal((vbits >> slot) & 1, sel.to(torch.uint8))
    keep = 255 & ~(1 << slot)
    assert torch.equal(vbits & keep, other & keep)  # other slots' bits kept
    assert int(mark.sum()) == 0
    assert torch.equal(packed[sel], full[sel])
    assert bool(torch.isnan(packed[~sel]).all())
    # inheriting: another slot takes slot's set plus the new marks
    mark2 = (torch.rand(T, generator=g, device="cuda") < 0.05).to(torch.uint8)
    sel2 = sel | (mark2 == 1)
    packed2 = torch.full_like(full, float("nan"))
    N.call("cmm_mwss_assemble_sparse", U.mesh._h, U.level_vals(1).data_ptr(), packed2.data_ptr(), vbits.data_ptr(),
           2, slot, mark2.data_ptr(), B.stream())
    torch.cuda.synchronize()
    assert torch.equal((vbits >> 2) & 1, sel2.to(torch.uint8))
    assert torch.equal(packed2[sel2], full[sel2])
    assert bool(torch.isnan(packed2[~sel2]).all())
    with pytest.raises(N.CmmError):
        N.call("cmm_mwss_assemble_sparse", U.mesh._h, U.level_vals(1).data_ptr(), packed2.data_ptr(),
               vbits.data_ptr(), 2, 2, mark2.data_ptr(), B.stream())
    vbits = (vbits & (255 & ~4)) | (other & 4)
    # completion: the rest, every bit of the slot set, the other blocks kept
    N.call("cmm_mwss_assemble_sparse", U.mesh._h, U.level_vals(1).data_ptr(), packed.data_ptr(), vbits.data_ptr(),
           slot, -1, None, B.stream())
    torch.cuda.synchronize()
    assert bool(((vbits >> slot) & 1 == 1).all())
    assert torch.equal(vbits & keep, other & keep)
    assert torch.equal(packed, full)


def test_loop_with_sparse_levels_matches_full_levels():
    """BarotropicCMM with the velocity assembled where the feet fall against
    the same run with complete levels: identical maps and vorticity; the
    assembled fraction stays well below one."""
    from cmm_s2.core import dynamical_fields as df, mesh_functions as mf, spherical_harmonic_tools as st
    from cmm_s2.core.spherical_simulations import BarotropicCMM
    L, Nt = 48, 30
    ico = mf.icosahedral_mesh(4)
    mesh = mf.spherical_triangulation(ico.points, simplices=ico.simplices)
    w0 = st.project_onto_S12(df.lm_to_full(df.random_sh_coefficients(11, 12, L), L), L,
                             mf.structure_spherical_triangulation(L))
    runs = {}
    for sp in (False, True):
        sim = BarotropicCMM(L, Nt, Nt / L, mesh, w0, 2 * np.pi, n_maps=10, sparse_velocity=sp)
        sim.run()
        torch.cuda.synchronize()
        frac = max(sim.U.valid_fraction(i) for i in range(3))
        runs[sp] = (sim.submaps.packed[:sim.submaps.Nc].clone(), sim.vorticity_on_grid(2 * L).clone(), frac)
    assert torch.equal(runs[True][0], runs[False][0])
    assert torch.equal(runs[True][1], runs[False][1])
    assert runs[False][2] == 1.0
    assert runs[True][2] < 0.8, runs[True][2]


def test_adjoint_entry_points_refuse_sparse_levels():
    from cmm_s2 import _native as N
    from cmm_s2.core import mesh_functions as mf, evolution_functions as ev
    from cmm_s2.core import _backend as B
    U = _velocity(16, 5)
    V = _sparse_copy(U, 0.5, 1)
    ico = mf.icosahedral_mesh(2)
    mesh = mf.spherical_triangulation(ico.points, simplices=ico.simplices)
    xs = mesh.d_spts
    desc = ev.velocity_descriptor(V, ev._STAGES[N.INT_RK4](0.1, 0.05), 0.05)
    yb = torch.zeros_like(xs)
    bars = [torch.zeros_like(V.level_packed(k)) for k in range(3)]
    with pytest.raises(N.CmmError, match="sparse"):
        N.call("cmm_advect_stencil_adjoint", desc, N.INT_RK4, 0.05, xs.data_ptr(), xs.stride(0), mesh.nv,
               yb.data_ptr(), yb.stride(0), N.ptr_array([b.data_ptr() for b in bars]), None, 0, None, 0,
               B.stream())
