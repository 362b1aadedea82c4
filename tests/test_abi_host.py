"""CPU: the C-ABI library loads and exports every entry point include/cmm_s2.h
declares, argument errors are reported without touching the GPU, and the
host-side mesh / spectral set-up of the product matches the oracle."""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    txt = open(os.path.join(ROOT, "include", "cmm_s2.h")).read()
    return sorted(set(re.findall(r"^(?:int|const char\*)\s+(cmm_[a-z0-9_]+)\(", txt, re.M)))


def test_library_exports_every_declared_symbol():
    from cmm_s2 import _native as N
    lib = N.lib()
    names = _declared()
    assert len(names) >= 27
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(N.SIGNATURES)
    assert lib.cmm_abi_version() == 1


def test_argument_errors_without_gpu():
    from cmm_s2 import _native as N
    lib = N.lib()
    h = ctypes.c_void_p()
    assert lib.cmm_tri_mesh_create(None, 0, None, 0, 0, ctypes.byref(h)) == -1
    assert b"bad arguments" in lib.cmm_last_error()
    v = np.zeros((3, 3))
    s = np.array([[0, 1, 7]], dtype=np.int64)
    assert lib.cmm_tri_mesh_create(N.dptr(v), 3, N.lptr(s), 1, 0, ctypes.byref(h)) == -1
    assert b"out of range" in lib.cmm_last_error()
    th = np.zeros(3)
    assert lib.cmm_mwss_mesh_create(1, N.dptr(th), N.dptr(th), N.dptr(th), N.dptr(th), N.dptr(th),
                                    N.dptr(th), ctypes.byref(h)) == -1
    ct = np.zeros(4098)
    # any L up to 2048 (ring FFTs of any length 2L); beyond it the two LDS
    # buffers of a ring FFT no longer fit
    assert lib.cmm_sht_plan_create(2049, N.dptr(ct), N.dptr(ct), ctypes.byref(h)) == -3
    assert lib.cmm_sht_plan_create(4097, N.dptr(ct), N.dptr(ct), ctypes.byref(h)) == -3
    vel = N.Velocity()
    vel.kind = 9
    assert lib.cmm_advect_points(ctypes.byref(vel), 2, 0.1, None, 0, None, None) == -1
    # grouped-scatter sizes and the sparse velocity adjoint: checked before
    # any device work
    nb = ctypes.c_int64(0)
    assert lib.cmm_scatter_workspace_bytes(0, 10, ctypes.byref(nb)) == -1
    assert lib.cmm_scatter_workspace_bytes(1 << 31, 10, ctypes.byref(nb)) == -3
    assert lib.cmm_scatter_workspace_bytes(10, 1 << 32, ctypes.byref(nb)) == -3
    # hipCUB counts key_range + 1 bins in an int: 2^31 - 1 keys would wrap it
    assert lib.cmm_scatter_workspace_bytes(10, (1 << 31) - 1, ctypes.byref(nb)) == -3
    vel.kind = N.VEL_SPLINE
    assert lib.cmm_advect_stencil_adjoint_corners(ctypes.byref(vel), N.INT_RK4, 0.1, None, 0, 0, None, 0, None,
                                                  None, None, 0, None, 0, None) == -1
    assert b"cmm_advect_stencil_adjoint_corners" in lib.cmm_last_error()
    assert lib.cmm_mwss_corner_jets_adjoint(None, None, None, None, None) == -1
    with pytest.raises(N.CmmError):
        N.check(-1, "x")


@pytest.mark.parametrize("k", [0, 2, 4])
def test_icosahedral_generator(k):
    from cmm_s2.core import _meshgen as G
    from oracle import meshes as om
    ico = G.IcosahedralMesh(k)
    P, S = ico.points, ico.simplices
    assert len(P) == 10 * 4 ** k + 2 and len(S) == 20 * 4 ** k
    assert np.allclose(P[0], [0, 0, 1]) and np.allclose(P[-1], [0, 0, -1])
    v = P[S]
    det = np.einsum("ij,ij->i", v[:, 0], np.cross(v[:, 1], v[:, 2]))
    assert (det > 0).all()
    # every edge shared by exactly two triangles (closed surface), V - E + F = 2
    e = np.sort(np.concatenate([S[:, [0, 1]], S[:, [1, 2]], S[:, [2, 0]]]), axis=1)
    u, c = np.unique(e, axis=0, return_counts=True)
    assert (c == 2).all() and len(P) - len(u) + len(S) == 2
    o = om.icosahedral_points(k)
    assert np.allclose(np.sort(o, axis=0), np.sort(P, axis=0), atol=1e-15)


def test_hull_triangulation_and_refinement():
    from cmm_s2.core import _meshgen as G
    ico = G.IcosahedralMesh(1)
    S = G.hull_triangles(ico.points)
    key = lambda T: {tuple(sorted(t)) for t in T.tolist()}
    assert key(S) == key(ico.simplices)
    V2, S2 = G.refine(ico.points, ico.simplices, 2)
    ico3 = G.IcosahedralMesh(3)
    assert len(V2) == len(ico3.points) and len(S2) == len(ico3.simplices)


def test_product_stencil_points_match_oracle():
    from cmm_s2.core import mesh_functions as mf
    from oracle import meshes as om
    P = mf.icosahedral_mesh(3).points
    s1, t1 = mf.stencil_points(P)
    s2, t2 = om.stencil_points(P)
    assert np.array_equal(s1, s2) and np.array_equal(t1[0], t2[0]) and np.array_equal(t1[1], t2[1])


def test_product_ps_split_tables_match_oracle():
    from cmm_s2.core import mesh_functions as mf
    from oracle import meshes as om
    ico = mf.icosahedral_mesh(2)
    a = mf.ps_split_tables(ico.points, ico.simplices)
    b = om.ps_split(ico.points, ico.simplices)
    for x, y in zip(a, b):
        for u, w in zip(x, y):
            assert np.abs(np.asarray(u) - np.asarray(w)).max() < 1e-15


def test_spectral_host_helpers():
    from cmm_s2.core import spherical_harmonic_tools as st, spherical_spline as ss
    from oracle import sht as osh, splines as osp
    L = 9
    a = np.arange(L * L) + 1j * np.arange(L * L)[::-1]
    assert np.array_equal(st.coeff_array(a, L), osh.coeff_array(a, L))
    assert np.array_equal(st.lm_array(st.coeff_array(a, L), L), a)
    for i, t in enumerate(st.operator_tables(L)):
        assert np.array_equal(t, osh.operator_tables(L)[i])
    assert st.nlm(L) == L * (L + 1) // 2 and st.lm_index(L, L - 1, L - 1) == st.nlm(L) - 1
    for nVs in (1, 2, 3):
        assert ss.lagrange_weights(nVs, 0.27, 0.1, 0.05) == osp.lagrange_weights(nVs, 0.27, 0.1, 0.05)
    assert np.array_equal(ss.Cs, osp.Cs) and np.array_equal(ss.edges_ps, osp.edges_ps)


def test_random_field_generator_matches_oracle():
    from cmm_s2.core import dynamical_fields as df
    from oracle import fields as ofl
    assert np.array_equal(df.random_sh_coefficients(4, 32, 64), ofl.random_sh_coefficients(4, 32, 64))
    full = df.lm_to_full(ofl.random_sh_coefficients(4, 8, 16), 16)
    from oracle import sht as osh
    assert np.array_equal(full, osh.lm_to_full(ofl.random_sh_coefficients(4, 8, 16), 16))


def test_checkpoint_rejects_foreign_files(tmp_path):
    import torch
    from safetensors.torch import save_file
    from cmm_s2 import _native as N
    from cmm_s2.core import checkpoint as ck
    p = str(tmp_path / "x.safetensors")
    save_file({"a": torch.zeros(2)}, p, metadata={"cmm": '{"format": "something-else"}'})
    with pytest.raises(N.CmmError):
        ck.read_state(p)
