"""Analytic plugin fields (oracle).  Restates cmm_s2/core/dynamical_fields.py
plus the two workload inputs that are not in the reference:

* ``u_rh4`` -- Rossby-Haurwitz wavenumber-4 velocity u = x cross grad(psi),
  psi = -w z + K z Re(e^{-4 i nu t} (x+iy)^4), i.e. the Haurwitz stream
  function -w sin(lat) + K cos^4(lat) sin(lat) cos(4(lon - nu t)) extended as
  a polynomial; nu = (R(3+R)w - 2 Omega)/((1+R)(2+R)), R = 4.
* ``random_sh_coefficients`` -- seeded Gaussian spherical-harmonic field with
  variance l^-3 for 1 <= l <= lmax.

Pinned by tests/test_oracle_fields.py: u_rh4 tangency and agreement with a
central-difference gradient of psi; u_sbr is the exact solid-body rotation;
rossby_wave equals the closed-form RH(5,4) vorticity of the reference test.
"""
import numpy as np

from . import geometry as g

pi = np.pi


def u_sbr(t, dt, X, T=1.0, alpha=0.0):
    """dynamical_fields.py:18-22."""
    c = -2 * pi / T
    return np.array(g.cross(X, [c * np.sin(alpha), 0, c * np.cos(alpha)]))


def u_deform(t, dt, X, T=1.0):
    """dynamical_fields.py:24-30."""
    k = 2
    c1 = 2 * k * np.cos(pi * t / T)
    return np.array(g.cross(X, [0 * X[0], c1 * X[1], 0 * X[2]]))


def _Rz(a):
    return np.array([[np.cos(a), -np.sin(a), 0], [np.sin(a), np.cos(a), 0], [0, 0, 1]])


def _Ry(a):
    return np.array([[np.cos(a), 0, -np.sin(a)], [0, 1, 0], [np.sin(a), 0, np.cos(a)]])


def _rotate(M, X):
    return [M[0, 0] * X[0] + M[0, 1] * X[1] + M[0, 2] * X[2],
            M[1, 0] * X[0] + M[1, 1] * X[1] + M[1, 2] * X[2],
            M[2, 0] * X[0] + M[2, 1] * X[1] + M[2, 2] * X[2]]


def u_deform_rot(t, dt, X, alpha=0.0):
    """dynamical_fields.py:32-42."""
    tau = 2 * pi * t
    X_r = _rotate(np.matmul(_Rz(tau).T, _Ry(alpha)), X)
    u_s = u_deform(t, dt, X_r)
    u_sr = _rotate(np.matmul(_Ry(alpha).T, _Rz(tau)), u_s)
    return np.array(u_sr) + u_sbr(t, dt, X)


def cart2sphere(X):
    """utils.py:18-23: phi in [0, 2 pi), theta the colatitude."""
    phi = (np.arctan2(X[1], X[0]) + 2 * pi) % (2 * pi)
    theta = np.arctan2(np.sqrt(X[1] ** 2 + X[0] ** 2), X[2])
    return phi, theta


def _Rx(a):
    return np.array([[1, 0, 0], [0, np.cos(a), -np.sin(a)], [0, np.sin(a), np.cos(a)]])


RHO_0 = 3.0  # dynamical_fields.py:44


def omega_r(theta):
    """dynamical_fields.py:48-55: angular speed of the static vortex."""
    theta = np.asarray(theta, dtype=np.float64)
    s = np.sin(theta)
    out = np.zeros_like(theta)
    nz = theta != 0.0
    r = RHO_0 * s[nz]
    out[nz] = (2 * pi) * np.sqrt(3) * (3 / 2) * (1 / r) * np.tanh(r) * (1 / np.cosh(r)) ** 2
    return out


def u_static_vortex(t, dt, X):
    """dynamical_fields.py:57-66."""
    phi, theta = cart2sphere(X)
    omg = omega_r(theta)
    return np.array([-X[1] * omg, X[0] * omg, 0 * X[2]])


def u_rotating_vortex(t, dt, X, T=1.0):
    """dynamical_fields.py:68-78."""
    tau = 2 * pi * t / T
    X_r = _rotate(np.matmul(_Rx(pi / 2), _Rz(tau).T), X)
    u_p = u_static_vortex(t, dt, np.array(X_r))
    u_sr = _rotate(np.matmul(_Rz(tau), _Rx(pi / 2).T), u_p)
    return np.array(u_sr) + u_sbr(t, dt, X)


def equator_vortex(t, dt, X):
    """dynamical_fields.py:80-85."""
    X_r = _rotate(_Rx(pi / 2), X)
    u_p = u_static_vortex(t, dt, np.array(X_r))
    return np.array(_rotate(_Rx(pi / 2).T, u_p))


def u_div(t, dt, X, T=1.0):
    """dynamical_fields.py:99-108: (dphi/dt, dtheta/dt) = (u, v) mapped to
    Cartesian components, then projected to the tangent plane."""
    phi, theta = cart2sphere(X)
    u = -(np.sin(phi / 2) ** 2) * np.sin(2 * theta) * np.sin(theta) * np.cos(pi * t / T)
    v = (-1 / 2) * np.sin(phi) * (np.sin(theta) ** 3) * np.cos(pi * t / T)
    u_x = np.cos(theta) * np.cos(phi) * v - np.sin(theta) * np.sin(phi) * u
    u_y = np.cos(theta) * np.sin(phi) * v + np.sin(theta) * np.cos(phi) * u
    u_z = -np.sin(theta) * v
    return np.array(g.tan_proj(X, [u_x, u_y, u_z]))


def static_vort_IC(t, xyz):
    """dynamical_fields.py:87-90: the tracer the static vortex carries,
    exact at time t."""
    phi, theta = cart2sphere(xyz)
    rho = RHO_0 * np.sin(theta)
    return 1 - np.tanh((rho / 5) * np.sin(phi - omega_r(theta) * t))


def rotating_vort_IC(t, xyz):
    """dynamical_fields.py:92-97 (its frame angle tau is 0)."""
    X_r = _rotate(np.matmul(_Rx(pi / 2), _Rz(0.0).T), xyz)
    phi, theta = cart2sphere(np.array(X_r))
    rho = RHO_0 * np.sin(theta)
    return 1 - np.tanh((rho / 5) * np.sin(phi - omega_r(theta) * t))


# tracer initial conditions (dynamical_fields.py:110-156)
H_0 = 1.0
PHI_C1, THE_C1 = pi + pi / 6, pi / 2
PHI_C2, THE_C2 = pi - pi / 6, pi / 2
R_BELL = 1 / 2


def _bell_dist(phi, theta, phc, thc):
    arg = np.sin(thc - pi / 2) * np.sin(theta - pi / 2) + np.cos(thc - pi / 2) * np.cos(theta - pi / 2) * \
        np.cos(phi - phc)
    return np.arccos(arg)


def cosine_bell(phi, theta):
    """dynamical_fields.py:123-138."""
    out = np.zeros_like(np.asarray(theta, dtype=np.float64))
    out2 = np.zeros_like(out)
    r1 = _bell_dist(phi, theta, PHI_C1, THE_C1)
    r2 = _bell_dist(phi, theta, PHI_C2, THE_C2)
    i1, i2 = r1 < R_BELL, r2 < R_BELL
    out[i1] = (H_0 / 2) * (1 + np.cos(r1[i1] * pi / R_BELL))
    out2[i2] = (H_0 / 2) * (1 + np.cos(r2[i2] * pi / R_BELL))
    return 0.9 * (out + out2) + 0.1


def IC_SC(phi, theta):
    """dynamical_fields.py:140-156 (slotted cylinders)."""
    phi = np.asarray(phi, dtype=np.float64)
    theta = np.asarray(theta, dtype=np.float64)
    r1 = _bell_dist(phi, theta, PHI_C1, THE_C1)
    r2 = _bell_dist(phi, theta, PHI_C2, THE_C2)
    R = R_BELL
    out = np.zeros_like(theta)
    out[(r1 <= R) & (np.abs(phi - PHI_C1) >= R / 6)] = 0.9
    out[(r1 <= R) & (np.abs(phi - PHI_C1) < R / 6) & (theta < (-5 * R / 12 + pi / 2))] = 0.9
    out2 = np.zeros_like(theta)
    out2[(r2 <= R) & (np.abs(phi - PHI_C2) >= R / 6)] = 0.9
    out2[(r2 <= R) & (np.abs(phi - PHI_C2) < R / 6) & (theta > (5 * R / 12 + pi / 2))] = 0.9
    return out + out2 + 0.1


def random_ic_coeffs():
    """dynamical_fields.py:159-160: np.random.seed(42); uniform(-1, 1, 33^2)."""
    return np.random.RandomState(42).uniform(-1, 1, 33 ** 2)


def random_initial_condition(xyz):
    """dynamical_fields.py:161-171: sum over l <= 32 and m = -l .. l-1 (the
    reference's range stops before m = l) of c Re Y_l^m, coefficients in
    that order."""
    from scipy.special import sph_harm_y
    phi, theta = cart2sphere(xyz)
    c = random_ic_coeffs()
    out = 0.0
    lm = 0
    for l in range(33):
        for m in range(-l, l):
            out = out + c[lm] * sph_harm_y(l, m, theta, phi).real
            lm += 1
    return out


RH4_OMEGA = 7.848e-6
RH4_K = 7.848e-6
RH4_R = 4
ROT = 2 * pi


def rh4_nu(w=RH4_OMEGA, R=RH4_R, Omega=ROT):
    return (R * (3 + R) * w - 2 * Omega) / ((1 + R) * (2 + R))


def u_rh4(t, dt, X, w=RH4_OMEGA, K=RH4_K, Omega=ROT):
    """x cross grad psi for the RH-4 stream function (polynomial form)."""
    x, y, z = X[0], X[1], X[2]
    nu = rh4_nu(w, RH4_R, Omega)
    c, s = np.cos(4 * nu * t), np.sin(4 * nu * t)
    x2, y2 = x * x, y * y
    re4 = x2 * x2 - 6 * x2 * y2 + y2 * y2
    im4 = 4 * x2 * x * y - 4 * x * y2 * y
    A = c * re4 + s * im4
    dre_x, dre_y = 4 * x2 * x - 12 * x * y2, -12 * x2 * y + 4 * y2 * y
    dim_x, dim_y = 12 * x2 * y - 4 * y2 * y, 4 * x2 * x - 12 * x * y2
    px = K * z * (c * dre_x + s * dim_x)
    py = K * z * (c * dre_y + s * dim_y)
    pz = -w + K * A
    return np.array([y * pz - z * py, z * px - x * pz, x * py - y * px])


def psi_rh4(t, X, w=RH4_OMEGA, K=RH4_K, Omega=ROT):
    x, y, z = X[0], X[1], X[2]
    nu = rh4_nu(w, RH4_R, Omega)
    c, s = np.cos(4 * nu * t), np.sin(4 * nu * t)
    re4 = x ** 4 - 6 * x ** 2 * y ** 2 + y ** 4
    im4 = 4 * x ** 3 * y - 4 * x * y ** 3
    return -w * z + K * z * (c * re4 + s * im4)


# vorticities (dynamical_fields.py:176-214)
def rotating_frame(phi, theta):
    return 4 * pi * np.cos(theta)


def zonal_jet(phi, theta):
    beta2 = 12 ** 2
    theta_c = pi / 4
    u_lt = (pi / 2) * np.exp(-2 * beta2 * (1 - np.cos(-theta + pi / 2 - theta_c)))
    return (np.cos(-theta + pi / 2) * (2 * beta2 * (np.cos(theta_c) * np.sin(-theta + pi / 2)
            - np.sin(theta_c) * np.cos(-theta + pi / 2))) + np.sin(pi / 2 - theta)) * u_lt


def perturbed_zonal_jet(phi, theta):
    beta2 = 12 ** 2
    theta_c = pi / 4 + 0.01 * np.cos(12 * phi)
    u_lt = (pi / 2) * np.exp(-2 * beta2 * (1 - np.cos(-theta + pi / 2 - theta_c)))
    return (np.cos(-theta + pi / 2) * (2 * beta2 * (np.cos(theta_c) * np.sin(-theta + pi / 2)
            - np.sin(theta_c) * np.cos(-theta + pi / 2))) + np.sin(pi / 2 - theta)) * u_lt


def multi_jet(phi, theta):
    return perturbed_zonal_jet(phi, theta) + perturbed_zonal_jet(phi, theta - pi / 8)


def gaussian_vortex(phi, theta):
    x, y, z = g.sphere2cart(phi, theta)
    sigma, C = 1 / 16, 4 * pi
    return C * np.exp(-((x - 1) ** 2 + y ** 2 + z ** 2) / sigma)


def rossby_wave(phi, theta, t=0):
    ell = 5
    alpha = 1 / (ell * (ell + 1))
    Omega = 2 * pi
    return 30 * np.cos(theta) * np.cos(4 * (phi + 2 * Omega * alpha * t)) * np.sin(theta) ** 4


def perturbed_rossby_wave(phi, theta):
    return 30 * np.cos(theta) * np.cos(4 * phi) * np.sin(theta) ** 4 + 0.1 * np.sin(theta) * np.cos(30 * theta)


def random_sh_coefficients(seed, lmax, L):
    """[l, m>=0] complex array (L, L): a_l0 = s_l N, a_lm = s_l (N + iN)/sqrt2,
    s_l^2 = l^-3 for 1 <= l <= lmax, zero otherwise; draws in order l up,
    m up, real part then imaginary part."""
    rng = np.random.default_rng(seed)
    a = np.zeros((L, L), dtype=np.complex128)
    for l in range(1, lmax + 1):
        s = l ** -1.5
        a[l, 0] = s * rng.standard_normal()
        for m in range(1, l + 1):
            re = rng.standard_normal()
            im = rng.standard_normal()
            a[l, m] = complex(s * re / np.sqrt(2.0), s * im / np.sqrt(2.0))
    return a
