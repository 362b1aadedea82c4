"""GPU checks of the mixed-precision configuration: the barotropic run at
L = 1024 (1025 x 2048 grid) on the ico-7 map mesh (seed 3, 20 RK4 steps of
dt = 1/1024, remap every 10) with the Legendre synthesis in FP64, FP32,
split-BF16 with 5 and 6 digits and int8 digit tables with 5 and 6 digits.  Per output (psi_lm, u and grad u on the
grid, velocity coefficients, map positions at each remap, final vorticity)
the relative max-norm difference from the FP64 run, and the synthesis of
the same spectra against the FP64 oracle."""
import numpy as np
import pytest
import torch

from helpers import rel_err

pytestmark = pytest.mark.gpu

L, NT, NMAPS = 1024, 20, 10
# relative max-norm bounds per arithmetic: (against the FP64 run, synthesis
# against the FP64 oracle)
BOUNDS = {"fp64": (0.0, 1e-11), "fp32": (2e-6, 1e-6), "bf16x5": (2e-9, 2e-9), "bf16x6": (2e-10, 1e-11),
          "i8x5": (2e-9, 2e-9), "i8x6": (2e-10, 1e-11)}


@pytest.fixture(scope="module")
def runs():
    from cmm_s2.core import mesh_functions as mf, spherical_harmonic_tools as st, dynamical_fields as df
    from cmm_s2.core.spherical_simulations import BarotropicCMM
    print("config 3: L=1024 ico-7, fp64 / fp32 / bf16x5 / bf16x6 / i8x5 / i8x6", flush=True)
    ico = mf.icosahedral_mesh(7)
    mesh = mf.spherical_triangulation(ico.points, simplices=ico.simplices)
    coeffs = df.random_sh_coefficients(3, 64, L)
    w0 = st.project_onto_S12(df.lm_to_full(coeffs, L), L, mf.structure_spherical_triangulation(L))
    out = {}
    for mode in BOUNDS:
        sim = BarotropicCMM(L, NT, NT / L, mesh, w0, 2 * np.pi, NMAPS, legendre=mode)
        sim.run()
        out[mode] = (sim, sim.vorticity_on_grid(L).clone())
    torch.cuda.synchronize()
    return out


def _outputs(sim, w):
    sim.U.complete()
    dinv = torch.tensor([0.0] + [-1.0 / (l * (l + 1)) for l in range(1, L)], dtype=torch.float64,
                        device="cuda")[torch.from_numpy(sim.plan.l_of).cuda()]
    out = {"psi_lm": sim.w_lm[:, 0] * dinv[:, None], "u_grid": sim.fields[0:3], "grad_u_grid": sim.fields[3:12],
           "U_coefficients": sim.U.level_packed(2), "final_vorticity": w}
    for i in range(sim.submaps.Nc):
        out[f"map_positions_remap_{i}"] = sim.submaps.vals[i][:, 0]
    return out


@pytest.mark.parametrize("mode", ["fp32", "bf16x5", "bf16x6", "i8x5", "i8x6"])
def test_mixed_run_against_fp64_run(runs, mode):
    ref = _outputs(*runs["fp64"])
    got = _outputs(*runs[mode])
    assert runs[mode][0].submaps.Nc == 2
    errs = {k: float((got[k] - ref[k]).abs().max() / ref[k].abs().max()) for k in ref}
    print(mode, {k: f"{v:.2e}" for k, v in errs.items()}, flush=True)
    assert max(errs.values()) < BOUNDS[mode][0], errs
    # the forms are distinct arithmetic (not a silent FP64 path)
    assert errs["u_grid"] > 0.0


@pytest.fixture(scope="module")
def oracle_synthesis(runs):
    """the oracle's FP64 synthesis of the FP64 run's last velocity spectra,
    four of the 12 fields (u_x, u_y, u_z, one gradient row)"""
    from oracle import sht as osh
    sim = runs["fp64"][0]
    o = sim.ops.cpu().numpy()
    a = np.zeros((4, L, L), complex)
    a[:, sim.plan.l_of, sim.plan.m_of] = (o[:, :4, 0] + 1j * o[:, :4, 1]).T
    return osh.inverse_lm(a, L)


@pytest.mark.parametrize("mode", list(BOUNDS))
def test_synthesis_against_fp64_oracle(runs, oracle_synthesis, mode):
    """The FP64 run's last 12-field velocity spectra through each form's
    synthesis against the oracle's FP64 synthesis."""
    from cmm_s2.core import spherical_harmonic_tools as st
    got = st.SHT.get(L, mode).inverse(runs["fp64"][0].ops)[:4].cpu().numpy()
    e = rel_err(got, oracle_synthesis)
    print(mode, f"synthesis vs oracle {e:.2e}", flush=True)
    assert e < BOUNDS[mode][1]
