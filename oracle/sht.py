"""MWSS spherical-harmonic transforms and spectral operators (oracle).
Restates cmm_s2/core/spherical_harmonic_tools.py; the transforms replace
``pyssht.forward/inverse(..., Method="MWSS", Reality=True, backend="ducc")``
(call sites spherical_harmonic_tools.py:16,84,91,123,133), which are absent.

Conventions (ssht): orthonormal Y_lm with Condon-Shortley phase, bandlimit L
(0 <= l < L), MWSS grid theta_t = 2 t pi/(2L) (t = 0..L), phi_p = 2 p pi/(2L)
(p = 0..2L-1), real signals (f_{l,-m} = (-1)^m conj f_{lm}).

Transform algorithm (own, documented; identical definition on the GPU):
  inverse: F_m(theta_t) = sum_l f_lm Pbar_lm(cos theta_t) at the MWSS rings,
           then a real inverse DFT over the 2L longitudes.
  forward: F_m(theta_t) = (pi/L) * DFT over the ring; extend theta to the
           full circle with F_m(2pi - theta) = (-1)^m F_m(theta); resample that
           2L-periodic trigonometric polynomial (Nyquist term split as
           cos(L theta)) at the L Gauss-Legendre colatitudes; then
           f_lm = sum_j w_j Pbar_lm(x_j) G_m(theta_j).
  The forward is exact for band-limited input (pinned by round trips and
  Y_lm known answers); on non-band-limited input its aliasing is this
  algorithm's, and parity with ducc0 there is unpinned.
  Pbar_lm uses the standard three-term recurrence in l with a 2^600 scaled
  range (values whose scale stays below 2^-300 are flushed to zero).

Pinned by tests/test_oracle_sht.py: Pbar against scipy.special.sph_harm_y;
Gauss-Legendre nodes against numpy.polynomial.legendre.leggauss; synthesis
of single Y_lm against sph_harm_y on the grid; forward(inverse(a)) == a;
Rossby-Haurwitz (5,4) Poisson and velocity closed forms
(tests/test_spherical_harmonic_tools.py:48-61 of the reference);
L_z Y_lm = m Y_lm, L^2 = l(l+1) from the angular-momentum tables.
"""
import numpy as np

from . import geometry as g
from . import meshes
from . import splines as sp

SCALE_EXP = 600
_BIG = 2.0 ** SCALE_EXP
_HI = 2.0 ** 300
_LO = 2.0 ** -300


# ----------------------------------------------------------------------------
# Legendre functions, quadrature
# ----------------------------------------------------------------------------
def gauss_legendre(n):
    """Gauss-Legendre nodes (descending, x_j = cos theta_j with theta ascending)
    and weights by Newton iteration on P_n with asymptotic initial guesses."""
    j = np.arange(1, n + 1)
    x = np.cos(np.pi * (j - 0.25) / (n + 0.5))
    def pn(x):
        p0 = np.ones_like(x)
        p1 = x.copy()
        for k in range(2, n + 1):
            p0, p1 = p1, ((2 * k - 1) * x * p1 - (k - 1) * p0) / k
        return p1, n * (x * p1 - p0) / (x * x - 1)

    for _ in range(30):
        p1, dp = pn(x)
        dx = p1 / dp
        x = x - dx
        if np.max(np.abs(dx)) < 1e-15:
            break
    p1, dp = pn(x)
    x = x - p1 / dp
    p1, dp = pn(x)
    w = 2.0 / ((1 - x * x) * dp * dp)
    return x, w


def legendre_diag(L, sint):
    """Scaled Pbar_mm(x) for m = 0..L-1 at every node: returns (P, s) arrays of
    shape (L, nx); the value is P * 2^(600 s)."""
    nx = len(sint)
    P = np.zeros((L, nx))
    S = np.zeros((L, nx), dtype=np.int64)
    p = np.full(nx, 1.0 / np.sqrt(4.0 * np.pi))
    s = np.zeros(nx, dtype=np.int64)
    P[0], S[0] = p, s
    for m in range(1, L):
        p = p * (-np.sqrt((2.0 * m + 1.0) / (2.0 * m)) * sint)
        small = (np.abs(p) < _LO) & (p != 0)
        p = np.where(small, p * _BIG, p)
        s = np.where(small, s - 1, s)
        P[m], S[m] = p, s
    return P, S


def legendre_m(L, m, x, diag):
    """Pbar_lm(x) for l = m..L-1 at all nodes x -> (L-m, nx)."""
    Pd, Sd = diag
    out = np.zeros((L - m, len(x)))
    p0 = Pd[m].copy()
    s = Sd[m].copy()
    out[0] = np.where(s == 0, p0, 0.0)
    if m + 1 >= L:
        return out
    p1 = np.sqrt(2.0 * m + 3.0) * x * p0
    out[1] = np.where(s == 0, p1, 0.0)
    for l in range(m + 2, L):
        a = np.sqrt((4.0 * l * l - 1.0) / (l * l - m * m))
        b = np.sqrt(((l - 1.0) ** 2 - m * m) / (4.0 * (l - 1.0) ** 2 - 1.0))
        p2 = a * (x * p1 - b * p0)
        big = (s < 0) & (np.abs(p2) > _HI)
        p2 = np.where(big, p2 / _BIG, p2)
        p1 = np.where(big, p1 / _BIG, p1)
        s = np.where(big, s + 1, s)
        out[l - m] = np.where(s == 0, p2, 0.0)
        p0, p1 = p1, p2
    return out


def mwss_nodes(L):
    th = meshes.mwss_thetas(L)
    return th, np.cos(th), np.sin(th)


def _dirichlet(d, L):
    """(1/2L)[1 + 2 sum_{k=1}^{L-1} cos(k d) + cos(L d)]."""
    s = np.sin(d / 2)
    with np.errstate(divide="ignore", invalid="ignore"):
        val = np.sin((L - 0.5) * d) / s
    val = np.where(np.abs(s) < 1e-12, 2.0 * L - 1.0, val)
    return (val + np.cos(L * d)) / (2.0 * L)


def resample_matrices(L, theta_gl):
    """K_even, K_odd (L x (L+1)): G_m(theta_j) = sum_t K_{m%2}[j,t] F_m(theta_t)."""
    th = meshes.mwss_thetas(L)
    D1 = _dirichlet(theta_gl[:, None] - th[None, :], L)
    D2 = _dirichlet(theta_gl[:, None] + th[None, :], L)
    Ke = D1.copy()
    Ko = D1.copy()
    Ke[:, 1:L] += D2[:, 1:L]
    Ko[:, 1:L] -= D2[:, 1:L]
    return Ke, Ko


class SHTPlan:
    """Tables for one bandlimit L (all FP64)."""

    def __init__(self, L):
        self.L = L
        self.theta, self.x, self.sint = mwss_nodes(L)
        self.xgl, self.wgl = gauss_legendre(L)
        self.thgl = np.arccos(self.xgl)
        self.sgl = np.sqrt((1 - self.xgl) * (1 + self.xgl))
        self.Ke, self.Ko = resample_matrices(L, self.thgl)
        self.diag_mwss = legendre_diag(L, self.sint)
        self.diag_gl = legendre_diag(L, self.sgl)


_plans = {}


def plan(L):
    if L not in _plans:
        _plans[L] = SHTPlan(L)
    return _plans[L]


def inverse_lm(flm, L):
    """Synthesis.  flm: (nf, L, L) complex [l, m>=0] (entries m > l ignored).
    Returns real grids (nf, L+1, 2L)."""
    p = plan(L)
    flm = np.asarray(flm)
    single = flm.ndim == 2
    if single:
        flm = flm[None]
    nf = flm.shape[0]
    F = np.zeros((nf, L + 1, L + 1), dtype=np.complex128)
    for m in range(L):
        P = legendre_m(L, m, p.x, p.diag_mwss)          # (L-m, L+1)
        F[:, :, m] = np.einsum("lt,fl->ft", P, flm[:, m:, m])
    out = np.fft.irfft(F, n=2 * L, axis=2) * (2 * L)
    return out[0] if single else out


def forward_lm(f, L):
    """Analysis.  f: (nf, L+1, 2L) real.  Returns (nf, L, L) complex [l, m>=0]."""
    p = plan(L)
    f = np.asarray(f, dtype=np.float64)
    single = f.ndim == 2
    if single:
        f = f[None]
    nf = f.shape[0]
    F = np.fft.rfft(f, axis=2)[:, :, :L] * (np.pi / L)  # (nf, L+1, L)
    G = np.empty((nf, L, L), dtype=np.complex128)
    G[:, :, 0::2] = np.einsum("jt,ftm->fjm", p.Ke, F[:, :, 0::2])
    G[:, :, 1::2] = np.einsum("jt,ftm->fjm", p.Ko, F[:, :, 1::2])
    out = np.zeros((nf, L, L), dtype=np.complex128)
    for m in range(L):
        P = legendre_m(L, m, p.xgl, p.diag_gl)          # (L-m, L)
        out[:, m:, m] = np.einsum("lj,j,fj->fl", P, p.wgl, G[:, :, m])
    return out[0] if single else out


# ----------------------------------------------------------------------------
# coefficient layouts (spherical_harmonic_tools.py:30-47)
# ----------------------------------------------------------------------------
def lm_to_full(flm, L):
    """[l, m>=0] -> the reference's (L, 2L+1) array [l, l+m] (negative m by
    conjugate symmetry f_{l,-m} = (-1)^m conj f_lm; columns > 2l zero)."""
    out = np.zeros((L, 2 * L + 1), dtype=np.complex128)
    for l in range(L):
        for m in range(0, l + 1):
            out[l, l + m] = flm[l, m]
            if m > 0:
                out[l, l - m] = (-1) ** m * np.conj(flm[l, m])
    return out


def full_to_lm(c, L):
    """(L, 2L+1) [l, l+m] -> [l, m>=0] (the m >= 0 half the real transform reads)."""
    out = np.zeros((L, L), dtype=np.complex128)
    for l in range(L):
        out[l, :l + 1] = c[l, l:2 * l + 1]
    return out


def coeff_array(alms, L):
    """spherical_harmonic_tools.py:30-42 (pyssht index l^2 + l + m)."""
    outs = np.zeros([L, 2 * L + 1], dtype="complex128")
    for l in range(0, L):
        for m in range(0, 2 * l + 1):
            outs[l, m] = alms[l * l + l + (m - l)]
    return outs


def lm_array(coeffs, L):
    """spherical_harmonic_tools.py:44-47."""
    return np.array([coeffs[l, m] for l in range(0, L) for m in range(0, 2 * l + 1)])


def sht_forward(samples, L):
    """pysh.forward(samples, L, Method='MWSS', Reality=True) as the 1-D ssht
    array (index l^2+l+m)."""
    return lm_array(lm_to_full(forward_lm(samples, L), L), L)


def sht_inverse(flm1d, L):
    """pysh.inverse(flm, L, Method='MWSS', Reality=True): reads m >= 0."""
    return inverse_lm(full_to_lm(coeff_array(flm1d, L), L), L)


# ----------------------------------------------------------------------------
# spectral operators
# ----------------------------------------------------------------------------
def operator_tables(L):
    """spherical_simulations.py:195-205: D_inv, L_plus, L_minus, L_z."""
    els = np.arange(0, L)
    els[0] = 1
    D_inv = -1 / (els * (els + 1))
    D_inv[0] = 0.
    with np.errstate(invalid="ignore"):
        L_plus = np.fromfunction(lambda l, m: np.sqrt(l * (l + 1) - (m - l) * (m - l - 1)),
                                 (L, 2 * L + 1), dtype="float64")
        L_minus = np.fromfunction(lambda l, m: np.sqrt(l * (l + 1) - (m - l) * (m - l + 1)),
                                  (L, 2 * L + 1), dtype="float64")
    L_plus[np.where(np.isnan(L_plus))] = 0.
    L_minus[np.where(np.isnan(L_minus))] = 0.
    L_z = np.fromfunction(lambda l, m: m - l, (L, 2 * L + 1), dtype="float64")
    return D_inv, L_plus, L_minus, L_z


def inv_Laplacian(samples, L, L_inv):
    """spherical_harmonic_tools.py:14-19."""
    return lm_to_full(forward_lm(samples, L), L) * L_inv[:, None]


def angular_momentum(alms, L_plus, L_minus, L_z, temp):
    """spherical_harmonic_tools.py:49-62."""
    temp *= 0
    temp[:, 1::] = alms[:, 0:-1]
    out_plus = np.multiply(L_plus, temp)
    temp *= 0
    temp[:, 0:-1] = alms[:, 1::]
    out_minus = np.multiply(L_minus, temp)
    out_x = (out_plus + out_minus) / 2
    out_y = 1j * (out_minus - out_plus) / 2
    return [1j * out_x, 1j * out_y, 1j * np.multiply(L_z, alms)]


def proj(a, b):
    """spherical_harmonic_tools.py:145-148."""
    scl = (a[0] * b[0] + a[1] * b[1] + a[2] * b[2]) / np.sqrt(b[0] ** 2 + b[1] ** 2 + b[2] ** 2)
    return [scl * b[0], scl * b[1], scl * b[2]]


def tan_proj(U, R):
    """spherical_harmonic_tools.py:150-152."""
    pU = proj(U, R)
    return [U[0] - pU[0], U[1] - pU[1], U[2] - pU[2]]


def _inv_full(c, L):
    return inverse_lm(full_to_lm(c, L), L)


def project_stream(alms, L, L_plus, L_minus, L_z, U, n, temp):
    """spherical_harmonic_tools.py:66-105: psi_lm -> u = r x grad psi and
    grad u on the MWSS mesh vertices, written into U.vals[n] (J^T), then
    U.init_coeffs(n)."""
    if U.stepping:
        U.coeffs[0:2] = U.coeffs[1:3]
        U.vals[0:2] = U.vals[1:3]
    s_pts = U.mesh.grid_points
    u_lms = angular_momentum(alms, L_plus, L_minus, L_z, temp)
    for i in range(3):
        u_temp = _inv_full(u_lms[i], L)
        U.vals[n, i, 0, 1:-1] = u_temp[1:-1].reshape((L - 1) * 2 * L).real
        U.vals[n, i, 0, -1] = u_temp[-1, 0].real
        U.vals[n, i, 0, 0] = u_temp[0, 0].real
        flm_ang = angular_momentum(u_lms[i], L_plus, L_minus, L_z, temp)
        u_grad = np.array([_inv_full(flm_ang[k], L) for k in range(3)])
        u_grad[:] = tan_proj(g.cross(u_grad, s_pts), s_pts)
        U.vals[n, :, i + 1, 1:-1] = u_grad[:, 1:-1, :].reshape([3, (L - 1) * 2 * L]).real
        U.vals[n, :, i + 1, -1] = u_grad[:, -1, 0].real
        U.vals[n, :, i + 1, 0] = u_grad[:, 0, 0].real
    U.init_coeffs(n)


class ScalarSpline:
    """spherical_spline (spherical_spline.py:19-152) on the MWSS mesh."""

    def __init__(self, mesh, vals):
        self.mesh = mesh
        self.vals = vals
        self.coeffs = sp.assemble(vals, mesh.simplices, mesh.ps_split)

    def __call__(self, phi, the):
        return sp.scalar_eval(self.mesh.query, self.coeffs, phi, the)


def project_onto_S12(alms, L, mesh):
    """spherical_harmonic_tools.py:108-142."""
    vals = np.zeros([4, len(mesh.vertices)])
    _, L_plus, L_minus, L_z = operator_tables(L)
    temp = 0 * alms
    s_pts = mesh.grid_points
    u_lms = angular_momentum(alms, L_plus, L_minus, L_z, temp)
    u_grad = np.array([_inv_full(u_lms[i], L) for i in range(3)])
    u_grad[:] = tan_proj(g.cross(u_grad, s_pts), s_pts)
    vals[1::, 1:-1] = u_grad[:, 1:-1].reshape(3, (L - 1) * 2 * L).real
    vals[1::, -1] = u_grad[:, -1, 0].real
    vals[1::, 0] = u_grad[:, 0, 0].real
    vals_temp = _inv_full(alms, L)
    vals[0, 1:-1] = vals_temp[1:-1, :].reshape((L - 1) * 2 * L).real
    vals[0, -1] = vals_temp[-1, 0].real
    vals[0, 0] = vals_temp[0, 0].real
    return ScalarSpline(mesh, vals)


def energy_spectrum(alms, L):
    """spherical_harmonic_tools.py:157-163."""
    outs = [sum(alms[0, :] * alms[0, :].conjugate())]
    for ell in range(1, L):
        outs.append(sum(alms[ell, :] * alms[ell, :].conjugate()) / (ell * (ell + 1)))
    return np.array(outs)


def enstrophy(samples, L):
    """figures/data_gen_convergence.py:59-60, 88-89: |sum f_lm conj(f_lm)|
    over the full coefficient set of a real field (its m < 0 half by
    conjugate symmetry, as pysh.forward with Reality=False returns it)."""
    a = sht_forward(samples, L)
    return float(np.abs(np.sum(a * a.conjugate())))


def energy(samples, L):
    """figures/data_gen_convergence.py:61, 92: 0.5 |sum energy_spectrum|."""
    a = sht_forward(samples, L)
    return float(0.5 * np.abs(np.sum(energy_spectrum(coeff_array(a, L), L))))
