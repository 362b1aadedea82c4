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
