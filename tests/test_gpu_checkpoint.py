"""GPU: checkpoint / resume of the barotropic CMM loop (safetensors container
replacing the pickles of spherical_simulations.py:434-438, 453-454).  A run
saved mid-way and resumed in a fresh object ends bit-identical to an
uninterrupted run."""
import numpy as np
import pytest
import torch

from helpers import random_unit

pytestmark = pytest.mark.gpu


def _mesh():
    from cmm_s2.core import mesh_functions as mf
    ico = mf.icosahedral_mesh(2)
    return mf.spherical_triangulation(ico.points, simplices=ico.simplices)


def _spline_vorticity(L):
    from cmm_s2.core import dynamical_fields as df, mesh_functions as mf, spherical_harmonic_tools as st
    return st.project_onto_S12(df.lm_to_full(df.random_sh_coefficients(5, 6, L), L), L,
                               mf.structure_spherical_triangulation(L))


@pytest.mark.parametrize("field", ["rossby", "spline"])
def test_resume_is_bit_identical(tmp_path, field):
    from cmm_s2.core import dynamical_fields as df
    from cmm_s2.core.spherical_simulations import BarotropicCMM
    mesh = _mesh()
    w0 = df.rossby_wave if field == "rossby" else _spline_vorticity(16)
    L, Nt, T = 16, 14, 0.14
    ref = BarotropicCMM(L, Nt, T, mesh, w0, 2 * np.pi, n_maps=5)
    ref.run()
    part = BarotropicCMM(L, Nt, T, mesh, w0, 2 * np.pi, n_maps=5)
    part.bootstrap()
    while part.count < 7:
        part.step()
    path = str(tmp_path / "state.safetensors")
    part.save(path)
    # the file carries the initial vorticity: resume without passing it
    res = BarotropicCMM.resume(path, mesh)
    assert res.count == 7 and res.submaps.Nc == part.submaps.Nc == 1
    res.run()
    assert res.submaps.Nc == ref.submaps.Nc == 3
    assert torch.equal(res.submaps.vals[:3], ref.submaps.vals[:3])
    q = torch.from_numpy(random_unit(4000, 3)).cuda()
    assert torch.equal(res.submaps(q), ref.submaps(q))
    wg = res.vorticity_on_grid(24)
    assert torch.equal(wg, ref.vorticity_on_grid(24))


def test_run_writes_checkpoints(tmp_path):
    from cmm_s2.core import checkpoint as ck, dynamical_fields as df
    from cmm_s2.core.spherical_simulations import BarotropicCMM
    mesh = _mesh()
    path = str(tmp_path / "run.safetensors")
    sim = BarotropicCMM(16, 9, 0.09, mesh, df.rossby_wave, 2 * np.pi, n_maps=4)
    sim.run(checkpoint=path, every=3)
    tensors, meta = ck.read_state(path)
    assert meta["count"] == 9 and meta["Nc"] == 3 and meta["ns"] == [4, 8, 9]
    assert tuple(tensors["U_vals"].shape) == (3, 3, 4, sim.mesh_u.nv)
    other = _mesh()
    bad = type(mesh)(mesh.vertices[::-1].copy())
    with pytest.raises(Exception):
        BarotropicCMM.resume(path, bad)
    BarotropicCMM.resume(path, other)
