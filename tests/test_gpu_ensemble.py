"""GPU: an ensemble advanced on one HIP stream per member gives every member
the bits it has when run alone."""
import numpy as np
import pytest
import torch

from helpers import random_unit

pytestmark = pytest.mark.gpu


def test_members_on_streams_match_solo_runs():
    from cmm_s2.core import dynamical_fields as df, mesh_functions as mf
    from cmm_s2.core.ensemble import Ensemble
    from cmm_s2.core.spherical_simulations import BarotropicCMM
    meshes = {}
    for k in (2, 3):
        ico = mf.icosahedral_mesh(k)
        meshes[k] = mf.spherical_triangulation(ico.points, simplices=ico.simplices)
    specs = [(16, 2, 12, df.rossby_wave), (24, 3, 10, df.rossby_wave), (16, 2, 9, df.rossby_wave.at(t=0.3))]
    mk = lambda L, i, Nt, w: (lambda: BarotropicCMM(L, Nt, Nt / L, meshes[i], w, 2 * np.pi, n_maps=4))
    ens = Ensemble([mk(*s) for s in specs])
    ens.run()
    q = torch.from_numpy(random_unit(3000, 5)).cuda()
    for spec, m in zip(specs, ens.members):
        solo = mk(*spec)()
        solo.run()
        assert m.count == solo.count == spec[2] and m.submaps.Nc == solo.submaps.Nc
        assert torch.equal(m.submaps.vals[:m.submaps.Nc], solo.submaps.vals[:solo.submaps.Nc])
        assert torch.equal(m.submaps(q), solo.submaps(q))


def test_random_members_are_seeded():
    from cmm_s2.core.ensemble import random_members
    _, p1 = random_members(64)
    _, p2 = random_members(64)
    assert p1 == p2 and len(set(p1)) > 5
    assert {L for L, _ in p1} <= {128, 192, 256, 384, 512} and {i for _, i in p1} <= {5, 6, 7}


def _meshes():
    from cmm_s2.core import mesh_functions as mf
    out = {}
    for k in (2, 3):
        ico = mf.icosahedral_mesh(k)
        out[k] = mf.spherical_triangulation(ico.points, simplices=ico.simplices)
    return out


def _state(m):
    """Every buffer a run leaves results in (velocity levels completed: the
    loop assembles them where the feet fall)."""
    m.U.complete()
    nc = m.submaps.Nc
    return [m.submaps.packed[:nc].clone(), m.submaps.vals[:nc].clone(), m.curr.packed.clone(),
            m.U._packed.clone(), m.U._vals.clone(), m.X.clone(), m.w.clone(), m.fields.clone()]


def _clobber(m):
    for t in (m.submaps.packed, m.submaps.vals, m.curr.packed, m.U._packed, m.U._vals, m.X, m.w, m.w_lm,
              m.ops, m.fields):
        t.fill_(float("nan"))


@pytest.mark.parametrize("own_streams", [True, False])
def test_captured_run_replays_the_eager_bits(own_streams):
    """A full ensemble run recorded as one HIP graph: every replay (outputs
    overwritten with NaN before it, eager work between replays) writes the
    bits of the eager run."""
    from cmm_s2.core import dynamical_fields as df
    from cmm_s2.core.ensemble import Ensemble
    from cmm_s2.core.spherical_simulations import BarotropicCMM
    meshes = _meshes()
    specs = [(16, 2, 12, df.rossby_wave), (24, 3, 10, df.rossby_wave), (16, 2, 9, df.rossby_wave.at(t=0.3))]
    mk = lambda L, i, Nt, w: (lambda: BarotropicCMM(L, Nt, Nt / L, meshes[i], w, 2 * np.pi, n_maps=4))
    ens = Ensemble([mk(*s) for s in specs], streams=own_streams)
    ens.run()
    torch.cuda.synchronize()
    ref = [_state(m) for m in ens.members]
    ens.capture()
    for rep in range(3):
        for m in ens.members:
            _clobber(m)
        ens.replay()
        torch.cuda.synchronize()
        for k, (m, r) in enumerate(zip(ens.members, ref)):
            got = _state(m)
            for j, (a, b) in enumerate(zip(got, r)):
                assert torch.equal(a, b), (rep, k, j)
        other = mk(16, 2, 6, df.rossby_wave)()  # eager work between replays
        other.run()
        torch.cuda.synchronize()


def test_replay_with_new_initial_fields():
    """The initial vorticity spline is a graph input: rewriting its
    coefficients in place and replaying gives the run of the new field."""
    from cmm_s2.core import dynamical_fields as df, mesh_functions as mf, spherical_harmonic_tools as st
    from cmm_s2.core.ensemble import Ensemble
    from cmm_s2.core.spherical_simulations import BarotropicCMM
    meshes = _meshes()

    def field(seed, L):
        c = df.random_sh_coefficients(seed, 6, L)
        return st.project_onto_S12(df.lm_to_full(c, L), L, mf.structure_spherical_triangulation(L))

    specs = [(24, 3, 10), (32, 2, 9)]
    w_a = [field(40 + k, L) for k, (L, _, _) in enumerate(specs)]
    w_b = [field(60 + k, L) for k, (L, _, _) in enumerate(specs)]
    ens = Ensemble([(lambda L=L, i=i, Nt=Nt, w=w: BarotropicCMM(L, Nt, Nt / L, meshes[i], w, 2 * np.pi, n_maps=4))
                    for (L, i, Nt), w in zip(specs, w_a)])
    ens.capture()
    for w, wb in zip(w_a, w_b):
        w.packed.copy_(wb.packed)
    ens.replay()
    torch.cuda.synchronize()
    for (L, i, Nt), m, wb in zip(specs, ens.members, w_b):
        solo = BarotropicCMM(L, Nt, Nt / L, meshes[i], wb, 2 * np.pi, n_maps=4)
        solo.run()
        assert m.submaps.Nc == solo.submaps.Nc == 3
        assert torch.equal(m.submaps.packed, solo.submaps.packed)
        assert torch.equal(m.curr.packed, solo.curr.packed)
        assert torch.equal(m.X, solo.X)


def test_batched_steps_match_solo_runs():
    """Every step issued as one launch per kernel for all members (different
    bandlimits, map meshes, run lengths and a second Legendre arithmetic):
    each member keeps the bits of its own run, and a step of members of one
    kind issues as many grids as one member's step has kernels."""
    from cmm_s2.core import dynamical_fields as df
    from cmm_s2.core.ensemble import Ensemble
    from cmm_s2.core.spherical_simulations import BarotropicCMM
    meshes = _meshes()
    specs = [(16, 2, 12, df.rossby_wave, "fp64"), (24, 3, 10, df.rossby_wave, "fp64"),
             (16, 2, 9, df.rossby_wave.at(t=0.3), "fp64"), (32, 3, 12, df.rossby_wave, "fp32"),
             (24, 2, 12, df.rossby_wave.at(t=0.1), "i8x5")]
    mk = lambda L, i, Nt, w, lg: (lambda: BarotropicCMM(L, Nt, Nt / L, meshes[i], w, 2 * np.pi, n_maps=4,
                                                        legendre=lg))
    # kernels of one member's steady step
    one = Ensemble([mk(*specs[0])])
    one.bootstrap()
    one.batched_step()
    per_step = one.batched_launches
    torch.cuda.synchronize()
    one.close()
    assert 10 <= per_step <= 16, per_step
    ens = Ensemble([mk(*s) for s in specs[:3]] + [mk(*s) for s in specs[3:]])
    ens.bootstrap()
    torch.cuda.synchronize()
    launches = []
    while True:
        k = ens.batched_step()
        if not k:
            break
        launches.append((k, ens.batched_launches))
    torch.cuda.synchronize()
    # first steady step: three kinds (fp64 / fp32 / int8 synthesis)
    assert launches[0] == (5, 3 * per_step + 1), launches[0]
    q = torch.from_numpy(random_unit(3000, 6)).cuda()
    for spec, m in zip(specs, ens.members):
        solo = mk(*spec)()
        solo.run()
        torch.cuda.synchronize()
        assert m.count == solo.count == spec[2] and m.submaps.Nc == solo.submaps.Nc
        for j, (a, b) in enumerate(zip(_state(m), _state(solo))):
            assert torch.equal(a, b), (spec[:3], j)
        assert torch.equal(m.submaps(q), solo.submaps(q))
    ens.close()


def test_partitioned_batched_steps_match_solo_runs():
    """Members split into two partitions, each batched on its own stream:
    the partitions cover every member once and members keep their bits."""
    from cmm_s2.core import dynamical_fields as df
    from cmm_s2.core.ensemble import Ensemble
    from cmm_s2.core.spherical_simulations import BarotropicCMM
    meshes = _meshes()
    specs = [(16, 2, 9), (24, 3, 10), (32, 3, 8), (16, 2, 11)]
    mk = lambda L, i, Nt: (lambda: BarotropicCMM(L, Nt, Nt / L, meshes[i], df.rossby_wave, 2 * np.pi, n_maps=4))
    ens = Ensemble([mk(*s) for s in specs])
    ens.batched_run(partitions=2)
    assert sorted(i for p in ens.partitions for i in p) == list(range(len(specs)))
    assert len(ens.partitions) == 2
    for spec, m in zip(specs, ens.members):
        solo = mk(*spec)()
        solo.run()
        torch.cuda.synchronize()
        assert m.count == solo.count == spec[2]
        for j, (a, b) in enumerate(zip(_state(m), _state(solo))):
            assert torch.equal(a, b), (spec, j)
    ens.close()


def test_batched_steps_across_block_size_classes():
    """Members above bandlimit 512 run some kernels with other block sizes
    (forward Legendre waves, resampling split): batched_step groups them
    apart from the small ones, and every member keeps its solo bits."""
    from cmm_s2.core import dynamical_fields as df
    from cmm_s2.core.ensemble import Ensemble
    from cmm_s2.core.spherical_simulations import BarotropicCMM
    meshes = _meshes()
    specs = [(24, 2, 5), (520, 2, 4), (32, 2, 5)]
    mk = lambda L, i, Nt: (lambda: BarotropicCMM(L, Nt, Nt / L, meshes[i], df.rossby_wave, 2 * np.pi, n_maps=3))
    ens = Ensemble([mk(*s) for s in specs])
    ens.bootstrap()
    torch.cuda.synchronize()
    groups = []
    while True:
        k = ens.batched_step()
        if not k:
            break
        groups.append((k, ens.batched_launches))
    torch.cuda.synchronize()
    for spec, m in zip(specs, ens.members):
        solo = mk(*spec)()
        solo.run()
        torch.cuda.synchronize()
        assert m.count == solo.count == spec[2]
        for j, (a, b) in enumerate(zip(_state(m), _state(solo))):
            assert torch.equal(a, b), (spec, j)
    ens.close()


def test_batch_recording_refuses_unbatched_entry_points():
    """While a batch records, an entry point without a batched form returns
    CMM_E_STATE instead of launching out of order; members that diverge are
    reported by cmm_batch_launch."""
    import ctypes
    from cmm_s2 import _native as N
    from cmm_s2.core import mesh_functions as mf
    h = N.vp()
    N.call("cmm_batch_create", 1 << 20, ctypes.byref(h))
    b = h.value
    mesh = _meshes()[2]
    q = torch.from_numpy(random_unit(100, 1)).cuda()
    tri = torch.empty(100, dtype=torch.int32, device="cuda")
    bcc = torch.empty(100, 3, dtype=torch.float64, device="cuda")
    N.call("cmm_batch_begin", b)
    N.call("cmm_batch_member", b, 0)
    rc = N.lib().cmm_tri_locate(mesh._h, q.data_ptr(), 100, 3, 1, tri.data_ptr(), bcc.data_ptr(), None)
    assert rc == -4
    x = torch.zeros(64, dtype=torch.float64, device="cuda")
    y = torch.ones(64, dtype=torch.float64, device="cuda")
    N.call("cmm_copy_f64", x.data_ptr(), y.data_ptr(), 64, None)
    N.call("cmm_batch_member", b, 1)  # a member with no launch diverges
    n = ctypes.c_int64()
    assert N.lib().cmm_batch_launch(b, None, ctypes.byref(n)) == -4
    torch.cuda.synchronize()
    assert float(x.abs().max()) == 0.0  # nothing ran
    # the same copy batched for two members runs both
    x2 = torch.zeros(64, dtype=torch.float64, device="cuda")
    N.call("cmm_batch_begin", b)
    for k, d in enumerate((x, x2)):
        N.call("cmm_batch_member", b, k)
        N.call("cmm_copy_f64", d.data_ptr(), y.data_ptr(), 64, None)
    N.call("cmm_batch_launch", b, None, ctypes.byref(n))
    torch.cuda.synchronize()
    assert n.value == 1 and torch.equal(x, y) and torch.equal(x2, y)
    N.call("cmm_batch_destroy", b)
