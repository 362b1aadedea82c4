"""Analytic plugin fields (cmm_s2/core/dynamical_fields.py API).

Velocity fields keep the signature V(t, dt, X(3,N)) -> (3,N) and vorticities
w(phi, theta) -> array; each object carries the id of its HIP
implementation, so the integrators and the vorticity sampler evaluate it
inside the fused kernels.  Called directly they also evaluate on the GPU.
"""
import numpy as np
import torch

from .. import _native as N
from . import _backend as B
from . import utils

pi = np.pi

RH4_OMEGA = 7.848e-6
RH4_K = 7.848e-6


class DeviceVelocity(object):
    def __init__(self, name, kind, params):
        self.__name__ = name
        self._cmm_kind = kind
        self._cmm_params = list(params)

    def __call__(self, t, dt, X):
        from .evolution_functions import integrate
        x = B.to_dev(X)
        return B.like(integrate(N.INT_EVAL, t, dt, x, self), X)


u_sbr = DeviceVelocity("u_sbr", N.VEL_SBR, [1.0, 0.0])                 # dynamical_fields.py:18-22
u_deform = DeviceVelocity("u_deform", N.VEL_DEFORM, [1.0])              # 24-30
u_deform_rot = DeviceVelocity("u_deform_rot", N.VEL_DEFORM_ROT, [1.0, 0.0])  # 32-42
u_static_vortex = DeviceVelocity("u_static_vortex", N.VEL_STATIC_VORTEX, [])      # 57-66
u_rotating_vortex = DeviceVelocity("u_rotating_vortex", N.VEL_ROTATING_VORTEX, [1.0])  # 68-78
equator_vortex = DeviceVelocity("equator_vortex", N.VEL_EQUATOR_VORTEX, [])      # 80-85
u_div = DeviceVelocity("u_div", N.VEL_DIV, [1.0])                       # 99-108
u_rh4 = DeviceVelocity("u_rh4", N.VEL_RH4, [RH4_OMEGA, RH4_K, 2 * pi])


def rh4_nu(w=RH4_OMEGA, R=4, Omega=2 * pi):
    """Haurwitz phase speed of the wavenumber-R wave."""
    return (R * (3 + R) * w - 2 * Omega) / ((1 + R) * (2 + R))


class DeviceVorticity(object):
    """An analytic field w(phi, theta) evaluated by a HIP kernel (kind id +
    4 parameters; a spherical-harmonic series also carries its coefficient
    table)."""

    def __init__(self, name, kind, params=(0.0, 0.0, 0.0, 0.0), table=None):
        self.__name__ = name
        self._cmm_kind = kind
        self._cmm_params = list(params) + [0.0] * (4 - len(params))
        self._table_host = table
        self._table = None

    @property
    def _cmm_table(self):
        """device coefficient table (series kinds), uploaded on first use"""
        if self._table_host is None:
            return None
        if self._table is None or self._table.device != B.device():
            self._table = B.to_dev(self._table_host)
        return self._table

    def at(self, **kw):
        """Copy with parameters (e.g. rossby_wave.at(t=1.0))."""
        p = list(self._cmm_params)
        if "t" in kw:
            p[0] = kw["t"]
        return DeviceVorticity(self.__name__, self._cmm_kind, p, self._table_host)

    def sample_xyz(self, X, out=None):
        """w0 at device points X (N,3) (no rotation term)."""
        n = X.shape[0]
        if out is None:
            out = B.empty(n)
        zero = B.zeros(1)
        N.call("cmm_sample_vorticity", self._cmm_kind, None, B.ptr(self._cmm_table), N.dbl_array(self._cmm_params),
               X.data_ptr(), n, X.stride(0), X.stride(1), zero.data_ptr(), n, 0.0, out.data_ptr(), B.stream())
        return out

    def __call__(self, phi, theta, t=None):
        f = self if t is None else self.at(t=t)
        p, th = B.to_dev(phi), B.to_dev(theta)
        X = torch.stack(utils.sphere2cart(p.reshape(-1), th.reshape(-1)), 1).contiguous()
        return B.like(f.sample_xyz(X).reshape(p.shape), phi)


rossby_wave = DeviceVorticity("rossby_wave", N.VORT_ROSSBY)          # dynamical_fields.py:205-210
zonal_jet = DeviceVorticity("zonal_jet", N.VORT_ZONAL_JET)            # 181-186
perturbed_zonal_jet = DeviceVorticity("perturbed_zonal_jet", N.VORT_PERTURBED_JET)  # 188-192
multi_jet = DeviceVorticity("multi_jet", N.VORT_MULTI_JET)            # 195-198
gaussian_vortex = DeviceVorticity("gaussian_vortex", N.VORT_GAUSSIAN) # 200-203
perturbed_rossby_wave = DeviceVorticity("perturbed_rossby_wave", N.VORT_PERTURBED_ROSSBY)  # 213-214


cosine_bell = DeviceVorticity("cosine_bell", N.VORT_COSINE_BELL)      # dynamical_fields.py:123-138
IC_SC = DeviceVorticity("IC_SC", N.VORT_SLOTTED_CYLINDER)             # 140-156


class DeviceTracer(object):
    """Tracer field f(t, xyz (3,N)) -> (N,) of dynamical_fields.py (exact
    solutions of the vortex flows), evaluated on the GPU; ``at(t)`` is the
    field at time t as a sampler of composed-map positions."""

    def __init__(self, name, kind):
        self.__name__ = name
        self._kind = kind

    def at(self, t):
        return DeviceVorticity(self.__name__, self._kind, [float(t)])

    def __call__(self, t, xyz):
        """xyz (3, N) as in the reference -> (N,)."""
        X = B.to_dev(xyz).reshape(3, -1).t().contiguous()
        return B.like(self.at(t).sample_xyz(X), xyz)


static_vort_IC = DeviceTracer("static_vort_IC", N.VORT_STATIC_VORTEX_TRACER)        # 87-90
rotating_vort_IC = DeviceTracer("rotating_vort_IC", N.VORT_ROTATING_VORTEX_TRACER)  # 92-97


def random_ic_table():
    """dynamical_fields.py:159-171 as real-field coefficients: np.random.seed(42)
    draws c = uniform(-1, 1, 33^2) consumed in the order l = 0..32,
    m = -l..l-1 (the reference's range stops before m = l) as
    f = sum c_lm Re Y_l^m; with Re Y_l^-m = (-1)^m Re Y_lm and
    Re Y_lm = (Y_lm + conj Y_lm) / 2 that is a_l0 = c_l0 and
    a_lm = (c_lm + (-1)^m c_l,-m) / 2 for m > 0 ([m][l] complex (33, 33))."""
    c = np.random.RandomState(42).uniform(-1, 1, 33 ** 2)
    d = np.zeros((33, 33))  # [l, m], coefficient of Re Y_lm (m >= 0)
    lm = 0
    for l in range(33):
        for m in range(-l, l):
            if m >= 0:
                d[l, m] += c[lm]
            else:
                d[l, -m] += (-1.0) ** (-m) * c[lm]
            lm += 1
    tab = np.zeros((33, 33, 2))
    for m in range(33):
        tab[m, :, 0] = d[:, m] * (1.0 if m == 0 else 0.5)
    return tab


class _SeriesField(DeviceVorticity):
    def __call__(self, xyz):
        """xyz (3, N) as in the reference -> (N,)."""
        X = B.to_dev(xyz).reshape(3, -1).t().contiguous()
        return B.like(self.sample_xyz(X), xyz)


random_initial_condition = _SeriesField("random_initial_condition", N.VORT_SH_SERIES, [33.0],
                                        table=random_ic_table())


def sh_series(a, name="sh_series"):
    """A real field from its coefficients a [l, m >= 0] complex (Ls, Ls)
    (the layout of random_sh_coefficients), evaluated exactly at any point."""
    a = np.asarray(a)
    Ls = a.shape[0]
    tab = np.zeros((Ls, Ls, 2))
    for m in range(Ls):
        tab[m, :, 0] = a[:, m].real
        tab[m, :, 1] = a[:, m].imag
    return DeviceVorticity(name, N.VORT_SH_SERIES, [float(Ls)], table=tab)


def random_sh_coefficients(seed, lmax, L):
    """Seeded Gaussian field: [l, m>=0] complex (L, L) with a_l0 = s_l N,
    a_lm = s_l (N + i N)/sqrt(2), s_l^2 = l^-3 for 1 <= l <= lmax (zero
    otherwise); draws in order l up, m up, real part then imaginary part."""
    rng = np.random.default_rng(seed)
    a = np.zeros((L, L), dtype=np.complex128)
    for l in range(1, lmax + 1):
        s = l ** -1.5
        a[l, 0] = s * rng.standard_normal()
        z = rng.standard_normal(2 * l).reshape(l, 2)
        a[l, 1:l + 1].real = s * z[:, 0] / np.sqrt(2.0)
        a[l, 1:l + 1].imag = s * z[:, 1] / np.sqrt(2.0)
    return a


def lm_to_full(a, L):
    """[l, m>=0] -> (L, 2L+1) [l, l+m] with conjugate symmetry."""
    out = np.zeros((L, 2 * L + 1), dtype=np.complex128)
    n = min(L, a.shape[0])
    for l in range(n):
        out[l, l:2 * l + 1] = a[l, :l + 1]
        m = np.arange(1, l + 1)
        out[l, l - m] = ((-1.0) ** m) * np.conj(a[l, 1:l + 1])
    return out
