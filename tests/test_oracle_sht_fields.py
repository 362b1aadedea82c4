"""CPU: pins of the oracle's spherical-harmonic transforms, spectral operators,
analytic fields and driver (spherical_harmonic_tools.py, dynamical_fields.py,
spherical_simulations.py)."""
import numpy as np
import pytest
from scipy.special import sph_harm_y

from helpers import random_unit
from oracle import fields as ofl, geometry as g, meshes as om, sht as osh, splines as osp


@pytest.mark.parametrize("L", [16, 40])
def test_legendre_against_scipy(L):
    p = osh.plan(L)
    for m in [0, 1, 7, L // 2, L - 1]:
        P = osh.legendre_m(L, m, p.x, p.diag_mwss)
        for l in sorted({m, m + 1, (m + L - 1) // 2, L - 1}):
            if l >= L or l < m:
                continue
            ref = sph_harm_y(l, m, p.theta, 0.0).real
            assert np.abs(P[l - m] - ref).max() < 1e-12 * max(1.0, np.abs(ref).max())


def test_legendre_scaled_range_at_large_order():
    """Near the poles Pbar_mm underflows double precision for large m; the
    scaled recurrence flushes values below 2^-300 and keeps the rest."""
    L = 512
    p = osh.plan(L)
    m = 300
    P = osh.legendre_m(L, m, p.x, p.diag_mwss)
    ref = sph_harm_y(np.arange(m, L)[:, None], m, p.theta[None, :], 0.0).real
    big = np.abs(ref) > 1e-250
    assert np.abs(P[big] - ref[big]).max() < 1e-11 * np.abs(ref).max()
    assert np.abs(P[~big]).max() < 1e-80


@pytest.mark.parametrize("n", [5, 64, 257])
def test_gauss_legendre_against_numpy(n):
    x, w = osh.gauss_legendre(n)
    xr, wr = np.polynomial.legendre.leggauss(n)
    assert np.abs(x - xr[::-1]).max() < 5e-15
    assert np.abs(w - wr[::-1]).max() < 5e-14 * max(1.0, wr.max())
    # exact for polynomials of degree <= 2n-1 (numpy's own nodes lose a few
    # digits at large n, so exactness is the sharper pin)
    for k in range(0, 2 * n, max(1, n // 4)):
        exact = 0.0 if k % 2 else 2.0 / (k + 1)
        assert abs(np.dot(w, x ** k) - exact) < 1e-14


@pytest.mark.parametrize("L", [8, 24])
def test_synthesis_of_single_harmonics(L):
    Phi, The = np.meshgrid(om.mwss_phis(L), om.mwss_thetas(L))
    for (l, m) in [(0, 0), (3, 0), (5, 2), (L - 1, L - 1), (L - 1, 1)]:
        a = np.zeros((L, L), complex)
        a[l, m] = 1.0 + 0.5j
        f = osh.inverse_lm(a, L)
        Y = sph_harm_y(l, m, The, Phi)
        ref = ((1.0 + 0.5j) * Y).real if m == 0 else 2 * ((1.0 + 0.5j) * Y).real
        assert np.abs(f - ref).max() < 1e-13


@pytest.mark.parametrize("L", [8, 32, 64])
def test_forward_inverts_synthesis(L):
    rng = np.random.default_rng(L)
    a = np.zeros((L, L), complex)
    for l in range(L):
        a[l, :l + 1] = rng.standard_normal(l + 1) + 1j * rng.standard_normal(l + 1)
    a[:, 0] = a[:, 0].real
    b = osh.forward_lm(osh.inverse_lm(a, L), L)
    assert np.abs(a - b).max() < 1e-12 * np.abs(a).max()
    # full reference layout round trip (ssht 1-D index l^2 + l + m)
    f = osh.inverse_lm(a, L)
    one_d = osh.sht_forward(f, L)
    assert np.abs(osh.coeff_array(one_d, L) - osh.lm_to_full(a, L)).max() < 1e-12 * np.abs(a).max()
    assert np.abs(osh.sht_inverse(one_d, L) - f).max() < 1e-12 * np.abs(f).max()


def test_poisson_and_velocity_closed_forms():
    """Rossby-Haurwitz (5,4): Delta psi = w with the closed forms of the
    reference's tests/test_spherical_harmonic_tools.py:48-61."""
    L = 32
    Phi, The = np.meshgrid(om.mwss_phis(L), om.mwss_thetas(L))
    D_inv, Lp, Lm, Lz = osh.operator_tables(L)
    psi = osh.inv_Laplacian(ofl.rossby_wave(Phi, The), L, D_inv)
    x, y, z = g.sphere2cart(Phi, The)
    sol = -z * ((1 - z ** 2) ** 2 - 8 * (1 - z ** 2) * x ** 2 + 8 * x ** 4)
    assert np.abs(osh._inv_full(psi, L) - sol).max() < 1e-12
    m = om.StructuredTriangulation(L)
    U = osp.VelocitySpline(m, np.zeros((3, 3, 4, len(m.vertices))), np.zeros((3, 3, 19, len(m.simplices))),
                           ts=np.zeros(3))
    osh.project_stream(psi, L, Lp, Lm, Lz, U, 0, 0 * psi)
    V = m.vertices
    xv, yv, zv = V[:, 0], V[:, 1], V[:, 2]
    psi_x = -32 * zv * xv ** 3 + 16 * zv * (1 - zv ** 2) * xv
    psi_z = -(1 - zv ** 2) ** 2 + 4 * zv ** 2 * (1 - zv ** 2) + 8 * xv ** 2 * (1 - 3 * zv ** 2) - 8 * xv ** 4
    vel = np.array([yv * psi_z, zv * psi_x - xv * psi_z, -yv * psi_x])
    assert np.abs(U.vals[0, :, 0, :] - vel).max() < 1e-11
    # gradients: tangential derivatives of the closed-form velocity by finite differences
    h = 1e-6
    for d in range(3):
        e = np.zeros(3)
        e[d] = 1.0
        Pe = e[None] - (V @ e)[:, None] * V
        Xp = V + h * Pe
        Xm = V - h * Pe
        def velf(X):
            x_, y_, z_ = X[:, 0], X[:, 1], X[:, 2]
            px = -32 * z_ * x_ ** 3 + 16 * z_ * (1 - z_ ** 2) * x_
            pz = -(1 - z_ ** 2) ** 2 + 4 * z_ ** 2 * (1 - z_ ** 2) + 8 * x_ ** 2 * (1 - 3 * z_ ** 2) - 8 * x_ ** 4
            return np.array([y_ * pz, z_ * px - x_ * pz, -y_ * px])
        fd = (velf(Xp) - velf(Xm)) / (2 * h)            # d/de of u_i, i = 0..2
        # U.vals[0, d, 1+i] = d_d u_i: the tangential derivative along P e_d
        assert np.abs(U.vals[0, d, 1:, :] - fd).max() < 1e-6 * np.abs(fd).max()


def test_angular_momentum_eigenvalues():
    L = 12
    _, Lp, Lm, Lz = osh.operator_tables(L)
    for (l, m) in [(3, 2), (7, -4), (11, 0)]:
        a = np.zeros((L, 2 * L + 1), complex)
        a[l, l + m] = 1.0
        ux, uy, uz = osh.angular_momentum(a, Lp, Lm, Lz, 0 * a)
        assert np.allclose(uz / 1j, m * a)                          # L_z Y = m Y
        # L^2 = -(iLx iLx + iLy iLy + iLz iLz) = l(l+1)
        L2 = -(osh.angular_momentum(ux, Lp, Lm, Lz, 0 * a)[0] + osh.angular_momentum(uy, Lp, Lm, Lz, 0 * a)[1]
               + osh.angular_momentum(uz, Lp, Lm, Lz, 0 * a)[2])
        assert np.allclose(L2, l * (l + 1) * a)


def test_energy_spectrum_of_single_modes():
    L = 8
    a = np.zeros((L, 2 * L + 1), complex)
    a[3, 3 + 2] = 2.0
    a[3, 3 - 2] = 2.0
    e = osh.energy_spectrum(a, L).real
    assert e[3] == pytest.approx(8.0 / 12.0) and np.abs(np.delete(e, 3)).max() == 0


def test_rh4_velocity_and_fields():
    X = random_unit(2000, 1).T
    u = ofl.u_rh4(0.3, 0.01, X)
    assert np.abs(np.einsum("ij,ij->j", u, X)).max() < 1e-20       # tangent
    h = 1e-5
    grad = np.zeros_like(X)
    for d in range(3):
        e = np.zeros((3, 1))
        e[d] = h
        grad[d] = (ofl.psi_rh4(0.3, X + e) - ofl.psi_rh4(0.3, X - e)) / (2 * h)
    ref = np.cross(X.T, grad.T).T
    assert np.abs(u - ref).max() < 1e-14
    # Haurwitz phase speed
    assert ofl.rh4_nu() == pytest.approx((4 * 7 * 7.848e-6 - 4 * np.pi) / 30)
    # solid-body rotation: u = x cross (0, 0, -2 pi)
    assert np.allclose(ofl.u_sbr(0, 0, X), np.cross(X.T, [0, 0, -2 * np.pi]).T)


def test_random_field_statistics_and_determinism():
    a = ofl.random_sh_coefficients(2, 64, 128)
    assert np.array_equal(a, ofl.random_sh_coefficients(2, 64, 128))
    assert np.abs(a[65:]).max() == 0 and a[0, 0] == 0
    ls = np.arange(1, 65)
    power = np.array([(np.abs(a[l, 0]) ** 2 + 2 * np.sum(np.abs(a[l, 1:l + 1]) ** 2)) / (2 * l + 1)
                      for l in ls])
    slope = np.polyfit(np.log(ls[8:]), np.log(power[8:]), 1)[0]
    assert -3.6 < slope < -2.4


@pytest.mark.slow
def test_driver_rossby_haurwitz_wave_travels():
    """RH(5,4) is an exact travelling solution of the barotropic vorticity
    equation (dynamical_fields.py:205-210; reference
    tests/test_barotropic_vorticity.py:40-43): after a short run the
    oracle's vorticity matches rossby_wave(t = T) up to the discretisation
    error, and the error drops with refinement."""
    from oracle import driver
    errs = []
    for k, L in [(2, 16), (3, 32)]:
        mesh = om.SphericalTriangulation(om.icosahedral_points(k))
        T, Nt = 0.04, 8
        sub = driver.euler_simulation_sphere_remapping(L, Nt, T, mesh, ofl.rossby_wave, 2 * np.pi, n_maps=5)
        q = random_unit(2000, 3)
        X = sub(q)
        ph, th = g.cart2sphere(X.T)
        w = ofl.rossby_wave(ph, th) + 4 * np.pi * np.cos(th) - 4 * np.pi * q[:, 2]
        qp, qt = g.cart2sphere(q.T)
        errs.append(np.abs(w - ofl.rossby_wave(qp, qt, t=T)).max())
    assert errs[1] < errs[0] / 4 and errs[1] < 2e-2, errs


def test_enstrophy_energy_closed_forms():
    """The oracle's diagnostics of figures/data_gen_convergence.py:58-95 on
    single spherical harmonics c Y_lm + conj: enstrophy = (1 or 2) |c|^2 over
    all orders, energy = half of it over l (l + 1) (the l = 0 term undivided),
    and both add over orthogonal modes."""
    from oracle import sht as osht
    L = 12
    tot_e = tot_n = 0.0
    flm_all = np.zeros((L, L), complex)
    for (l, m, c) in [(0, 0, 1.3), (3, 0, 0.7), (5, 2, 0.4 - 0.9j), (11, 11, 0.5j)]:
        flm = np.zeros((L, L), complex)
        flm[l, m] = c
        flm_all[l, m] = c
        f = osht.inverse_lm(flm, L)
        mult = 1 if m == 0 else 2
        e_true = mult * abs(c) ** 2
        n_true = 0.5 * e_true / (l * (l + 1) if l > 0 else 1)
        assert abs(osht.enstrophy(f, L) - e_true) < 1e-13 * e_true
        assert abs(osht.energy(f, L) - n_true) < 1e-13 * n_true
        tot_e += e_true
        tot_n += n_true
    f = osht.inverse_lm(flm_all, L)
    assert abs(osht.enstrophy(f, L) - tot_e) < 1e-13 * tot_e
    assert abs(osht.energy(f, L) - tot_n) < 1e-13 * tot_n


# ----------------------------------------------------------------------------
# forward transform on NON-band-limited input: the Nyquist term and the theta
# extension, pinned against the definition computed independently (numpy
# FFTs, 4L-node Gauss-Legendre quadrature, scipy's sph_harm_y) and against
# closed forms
# ----------------------------------------------------------------------------
def _lm_of(fun, L):
    from helpers import mwss_samples
    return osh.forward_lm(mwss_samples(fun, L), L)


def _projection_of_function(fun_theta, m, L, nodes):
    """2 pi int_0^pi fun_theta(theta) Pbar_lm(cos theta) sin theta dtheta for
    l = m..L-1, by Gauss-Legendre quadrature in cos theta (exact for the
    polynomial integrands used below)."""
    x, w = np.polynomial.legendre.leggauss(nodes)
    th = np.arccos(x)
    ls = np.arange(m, L)
    P = sph_harm_y(ls[:, None], m, th[None, :], 0.0).real
    return 2 * np.pi * (P * (w * fun_theta(th))[None, :]).sum(axis=1)


@pytest.mark.parametrize("L", [16, 24])
def test_forward_nyquist_term_closed_form(L):
    """cos(L theta) cos(m phi) (even m) samples to (-1)^t cos(m phi): pure
    Nyquist in theta.  Split as cos(L theta) it is reproduced exactly, so
    the transform must return the exact projection of cos(L theta) cos(m phi)
    onto l < L (T_L(x) Pbar_lm(x) has degree <= 2L - 1: quadrature exact).
    A dropped, doubled or one-sided Nyquist term fails."""
    for m in (0, 2, 6, L - 2):
        a = _lm_of(lambda T, P: np.cos(L * T) * np.cos(m * P), L)
        ref = _projection_of_function(lambda th: np.cos(L * th), m, L, 4 * L) * (1.0 if m == 0 else 0.5)
        assert m > 6 or np.abs(ref).max() > 0.05
        assert np.abs(a[m:, m] - ref).max() < 1e-13, m
        others = np.delete(a, m, axis=1)
        assert np.abs(others).max() < 1e-13


@pytest.mark.parametrize("L", [16, 24])
def test_forward_odd_order_extension_closed_form(L):
    """sin(k theta) cos(m phi) with odd m: the odd extension
    F_m(2 pi - theta) = -F_m(theta) reproduces sin(k theta) (k < L) exactly;
    for m >= 3 the field is not band-limited (it vanishes like theta, not
    theta^m, at the poles).  The transform must return its exact projection;
    an even extension for odd m fails."""
    for m in (1, 3, 5):
        for k in (L - 1, L // 2 + 1):
            a = _lm_of(lambda T, P: np.sin(k * T) * np.cos(m * P), L)
            ref = 0.5 * _projection_of_function(lambda th: np.sin(k * th), m, L, 4 * L)
            assert np.abs(ref).max() > 1e-3
            assert np.abs(a[m:, m] - ref).max() < 1e-13, (m, k)


def _smooth_fields():
    return {"exp": lambda T, P: np.exp(np.sin(T) * np.cos(P) + 0.5 * np.sin(T) * np.sin(P) - 0.3 * np.cos(T)),
            "bump": lambda T, P: 4 * np.pi * np.exp(-((np.sin(T) * np.cos(P) - 1) ** 2 + (np.sin(T) * np.sin(P)) ** 2
                                                     + np.cos(T) ** 2) / (1 / 16)),
            "cap": lambda T, P: np.tanh(8 * (np.cos(T) - 0.3)) * (1 + 0.2 * np.sin(T) * np.cos(3 * P))}


@pytest.mark.parametrize("L", [16, 24, 32])
def test_forward_non_band_limited_matches_definition(L):
    """Smooth fields that are not band-limited (functions on the sphere: pole
    rings constant): forward_lm equals the exact projection of the field's
    trigonometric interpolant with the split Nyquist term, computed
    independently; the same interpolant with a one-sided Nyquist term
    differs by far more than the tolerance (the check discriminates)."""
    from helpers import interpolant_projection, mwss_samples
    sens = {}
    for name, fun in _smooth_fields().items():
        f = mwss_samples(fun, L)
        a = osh.forward_lm(f, L)
        ref = interpolant_projection(f, L)
        scale = np.abs(ref).max()
        assert np.abs(a - ref).max() < 1e-12 * scale, name
        wrong = interpolant_projection(f, L, split_nyquist=False)
        sens[name] = np.abs(wrong - ref).max() / scale
    # the entire function "exp" has no Nyquist content at these L; the
    # steep fields do
    assert sens["cap"] > 1e-6 and sens["bump"] > 1e-9, sens


def test_theta_interpolant_split_matches_scipy_resample():
    """The split-Nyquist interpolant is the band-limited resampling of
    scipy.signal.resample (which splits an even-length signal's Nyquist
    component when upsampling): both evaluated on an 8L-point circle."""
    from scipy.signal import resample
    from helpers import mwss_samples, theta_interpolant
    L = 16
    f = mwss_samples(_smooth_fields()["cap"], L)
    th = 2 * np.pi * np.arange(8 * L) / (8 * L)
    G, s = theta_interpolant(f, L, th)
    R = resample(s, 8 * L, axis=0)
    assert np.abs(G - R).max() < 1e-13 * np.abs(R).max()
    G1, _ = theta_interpolant(f, L, th, split_nyquist=False)
    assert np.abs(G1 - R).max() > 1e-6 * np.abs(R).max()
