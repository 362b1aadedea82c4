"""GPU checks of the gradient configuration: J = sum_lm |(F omega_T -
omega_target)_lm|^2 of the barotropic run at L = 256 on the ico-6 map mesh
(50 RK4 steps of dt = 1/256, remap every 10, checkpoint at every remap),
gradient with respect to the initial coefficients 1 <= l <= 64 (4,224
reals), target seed 50, initial fields seeds 500..: central differences
along seeded unit directions, the same gradient with every segment
recomputed and with every segment tape kept, and a batch through the
autograd function."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def problem():
    from cmm_s2.core import mesh_functions as mf, spherical_harmonic_tools as st, dynamical_fields as df
    from cmm_s2.core.adjoint import MisfitProblem
    print("config 5: adjoint L=256 ico-6", flush=True)
    L, lmax = 256, 64
    im = mf.icosahedral_mesh(6)
    mesh = mf.spherical_triangulation(im.points, simplices=im.simplices)
    plan = st.SHT.get(L)
    target = plan.full_to_packed(df.lm_to_full(df.random_sh_coefficients(50, lmax, L), L))
    prob = MisfitProblem(L, 50, 50 / L, mesh, target, 2 * np.pi, 10, lmax=lmax)
    ps = [torch.from_numpy(prob.params.from_coeffs(df.random_sh_coefficients(500 + k, lmax, L))).cuda()
          for k in range(3)]
    return prob, ps


def test_gradient_against_central_differences(problem):
    """4 seeded unit directions, h = 1e-5: the truncation error is O(h^2)
    and the rounding noise ~1e-12 J / h (the stencil jets are 1/(4 eps)
    differences), both far below the 1e-4 bound."""
    prob, ps = problem
    assert prob.params.n == 4224
    J, g = prob.value_and_grad(ps[0])
    rng = np.random.default_rng(51)
    for i in range(4):
        d = torch.from_numpy(rng.standard_normal(prob.params.n)).cuda()
        d /= d.norm()
        ad = float((g * d).sum())
        h = 1e-5
        fd = (prob.value(ps[0] + h * d) - prob.value(ps[0] - h * d)) / (2 * h)
        print(f"direction {i}: adjoint {ad:.10e} fd {fd:.10e}", flush=True)
        assert abs(fd - ad) <= 1e-4 * abs(ad), (i, ad, fd)


def test_kept_and_recomputed_tapes_agree(problem):
    """The gradient from the default sweep (last segment's tape kept, the
    others recomputed on a second stream), from recomputing every segment
    on one stream, and from all segment tapes kept."""
    prob, ps = problem
    J0, g0 = prob.value_and_grad(ps[1])
    J1, g1 = prob.value_and_grad(ps[1], keep_bytes=0, overlap=False)
    J2, g2 = prob.value_and_grad(ps[1], keep_bytes=64 << 30)
    assert J0 == J1 == J2
    for g in (g1, g2):
        assert float((g - g0).norm()) <= 1e-12 * float(g0.norm())


def test_batch_through_autograd(problem):
    """misfit() on two fields summed into one loss before one backward():
    each field's gradient equals its own value_and_grad gradient."""
    from cmm_s2.core.adjoint import misfit
    prob, ps = problem
    xs = [p.clone().requires_grad_() for p in ps[:2]]
    loss = misfit(xs[0], prob) + misfit(xs[1], prob)
    loss.backward()
    for x, p in zip(xs, ps[:2]):
        _, g = prob.value_and_grad(p)
        assert float((x.grad - g).norm()) <= 1e-12 * float(g.norm())
