"""CPU pins of the plain-torch restatement of the CMM loop (tests/torch_cmm.py)
that supplies the autograd gradient the adjoint is checked against: its J
equals the NumPy oracle driver's (oracle/driver.py, the reference loop of
spherical_simulations.py:144-297) on the same inputs, and its autograd
gradient matches its own central differences."""
import functools

import numpy as np
import torch

import torch_cmm as T
from oracle import driver as odrv, fields as of, meshes as om, sht as osht
from oracle import geometry as og

L, NT, NMAPS, LMAX, ICO = 16, 8, 3, 4, 2


@functools.lru_cache(maxsize=1)
def _setup():
    pts = om.icosahedral_points(ICO)
    mesh = om.SphericalTriangulation(pts)
    tgt = of.random_sh_coefficients(50, LMAX, L)
    ref = T.TorchCMM(L, NT, NT / L, mesh.vertices, mesh.simplices, tgt, 2 * np.pi, NMAPS, lmax=LMAX)
    return mesh, tgt, ref


def _params(a):
    out = []
    for l in range(1, LMAX + 1):
        out.append(a[l, 0].real)
        for m in range(1, l + 1):
            out += [a[l, m].real, a[l, m].imag]
    return np.array(out)


def test_torch_loop_matches_oracle_driver():
    mesh, tgt, ref = _setup()
    a = of.random_sh_coefficients(500, LMAX, L)
    J = float(ref.J(torch.from_numpy(_params(a))))
    mu = om.StructuredTriangulation(L)
    w0 = osht.project_onto_S12(osht.lm_to_full(a, L), L, mu)
    sub = odrv.euler_simulation_sphere_remapping(L, NT, NT / L, mesh, w0, 2 * np.pi, NMAPS)
    q = mu.sample_points()
    X = sub(q)
    phi, the = og.cart2sphere(X.T if X.shape[0] != 3 else X)
    the_g = np.repeat(om.mwss_thetas(L), 2 * L)
    w = w0(phi.reshape(L + 1, 2 * L), the.reshape(L + 1, 2 * L)).reshape(-1) \
        + 4 * np.pi * np.cos(the) - 4 * np.pi * np.cos(the_g)
    c = osht.forward_lm(w.reshape(L + 1, 2 * L), L)
    d = c - tgt[:L, :L]
    cm = np.full((L, L), 2.0)
    cm[:, 0] = 1.0
    J_o = float((cm * np.tril(np.ones((L, L))) * np.abs(d) ** 2).sum())
    # the two loops differ by rounding only (Cramer vs back-substitution,
    # fused vs separate sums), amplified by the 1 / (4 eps) stencil jets
    assert abs(J - J_o) <= 1e-9 * J_o, (J, J_o)


def test_torch_autograd_matches_its_differences():
    _, _, ref = _setup()
    a = of.random_sh_coefficients(501, LMAX, L)
    p = torch.from_numpy(_params(a)).requires_grad_()
    J = ref.J(p)
    J.backward()
    g = p.grad.clone()
    rng = np.random.default_rng(3)
    for _ in range(2):
        d = torch.from_numpy(rng.standard_normal(p.numel()))
        d /= d.norm()
        h = 1e-5
        with torch.no_grad():
            fd = (float(ref.J(p + h * d)) - float(ref.J(p - h * d))) / (2 * h)
        ad = float((g * d).sum())
        assert abs(ad - fd) <= 1e-5 * abs(ad) + 2e-12 * float(J.detach()) / h, (ad, fd)
