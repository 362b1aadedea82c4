"""GPU check of the ensemble configuration: 64 members with seeded
bandlimits in {128, 192, 256, 384, 512} and icosahedral levels in {5, 6, 7}
(seed 4), l^-3 initial vorticity with l <= L/8 (seed 400 + k), 50 RK4 steps
of dt = 1/L with a remap every 10, FP32 synthesis, every member on its own
HIP stream: members of every (L, level) class give the bits of the same
member run alone."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def test_64_member_ensemble_matches_solo_runs():
    from cmm_s2.core.ensemble import Ensemble, random_members
    print("config 4: 64-member ensemble", flush=True)
    builders, params = random_members(64)
    assert len(params) == 64 and len(set(params)) >= 10
    ens = Ensemble(builders)
    ens.run()
    torch.cuda.synchronize()
    assert all(m.count == 50 and m.submaps.Nc == 5 for m in ens.members)
    # one member of each of the first four distinct classes, plus the last
    picks, seen = [], set()
    for k, p in enumerate(params):
        if p not in seen:
            seen.add(p)
            picks.append(k)
        if len(picks) == 4:
            break
    picks.append(63)
    q = None
    for k in picks:
        m = ens.members[k]
        solo = builders[k]()
        solo.run()
        torch.cuda.synchronize()
        nc = m.submaps.Nc
        assert torch.equal(m.submaps.packed[:nc], solo.submaps.packed[:nc]), k
        assert torch.equal(m.X, solo.X), k
        m.U.complete()
        solo.U.complete()
        assert torch.equal(m.U.level_packed(2), solo.U.level_packed(2)), k
        if q is None:
            g = torch.Generator().manual_seed(5)
            q = torch.randn(4096, 3, generator=g, dtype=torch.float64).cuda()
            q /= q.norm(dim=1, keepdim=True)
        assert torch.equal(m.submaps(q), solo.submaps(q)), k
        del solo
    ens.close()


def test_64_member_batched_steps_match_solo_runs():
    """The same 64 members with every step issued as one launch per kernel
    for all of them: members keep the bits of their solo runs."""
    from cmm_s2.core.ensemble import Ensemble, random_members
    print("config 4: 64-member ensemble, batched steps", flush=True)
    builders, params = random_members(64)
    ens = Ensemble(builders)
    ens.batched_run()
    assert all(m.count == 50 and m.submaps.Nc == 5 for m in ens.members)
    assert ens.batched_launches <= 16  # one partition: one launch per kernel
    for k in (0, 17, 63):
        m = ens.members[k]
        solo = builders[k]()
        solo.run()
        torch.cuda.synchronize()
        nc = m.submaps.Nc
        assert torch.equal(m.submaps.packed[:nc], solo.submaps.packed[:nc]), k
        assert torch.equal(m.X, solo.X), k
        del solo
    ens.close()
