"""GPU: cmm_sht_forward_noalloc (the analysis with its workspace in the
library's fixed device array) against cmm_sht_forward with a caller
workspace, through the C ABI: identical bits on two calls in a row, the
second with a smaller input (another bandlimit, fewer fields), and the
size limit reported before any device work."""
import ctypes

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _grid(L, nf, seed):
    g = torch.Generator().manual_seed(seed)
    f = torch.randn(nf, L + 1, 2 * L, generator=g, dtype=torch.float64)
    f[:, 0, :] = f[:, 0, :1]     # pole rings constant
    f[:, -1, :] = f[:, -1, :1]
    return f.cuda()


def _forward_with_workspace(plan, f, nf):
    from cmm_s2 import _native as N
    from cmm_s2.core import _backend as B
    from cmm_s2.core.spherical_harmonic_tools import nlm
    b = ctypes.c_int64(0)
    N.call("cmm_sht_workspace_bytes", plan._h, nf, ctypes.byref(b))
    work = torch.empty(b.value, dtype=torch.uint8, device="cuda")
    out = B.empty(nlm(plan.L), nf, 2)
    N.call("cmm_sht_forward", plan._h, f.data_ptr(), nf, out.data_ptr(), work.data_ptr(), B.stream())
    return out


def _forward_noalloc(plan, f, nf):
    from cmm_s2 import _native as N
    from cmm_s2.core import _backend as B
    from cmm_s2.core.spherical_harmonic_tools import nlm
    out = B.empty(nlm(plan.L), nf, 2)
    N.call("cmm_sht_forward_noalloc", plan._h, f.data_ptr(), nf, out.data_ptr(), B.stream())
    return out


def test_forward_noalloc_matches_workspace_form():
    from cmm_s2 import _native as N
    from cmm_s2.core.spherical_harmonic_tools import SHT
    for L, nf, seed in ((512, 3, 1), (128, 1, 2)):
        plan = SHT.get(L)
        f = _grid(L, nf, seed)
        ref = _forward_with_workspace(plan, f, nf)
        got = _forward_noalloc(plan, f, nf)
        torch.cuda.synchronize()
        assert torch.equal(got, ref), (L, nf)
        assert float(ref.abs().max()) > 0
    # beyond the 64 MiB scratch: refused on the host
    plan = SHT.get(512)
    f = _grid(512, 1, 3)
    out = torch.empty(1, dtype=torch.float64, device="cuda")
    # 16 fields at L = 512 need 32 * 16 * 512^2 = 134 MB > 64 MiB: refused by
    # the size check before any launch (f and out are never read or written)
    rc = N.lib().cmm_sht_forward_noalloc(plan._h, f.data_ptr(), 16, out.data_ptr(), None)
    assert rc == -3 and b"limit" in N.lib().cmm_last_error()
