"""Spectral tools (cmm_s2/core/spherical_harmonic_tools.py API) on the GPU.

The MWSS transforms that the reference takes from pyssht/ducc0 are the HIP
transforms of libcmm_s2_amd.so (ring FFTs in LDS, Legendre contractions with
on-the-fly recurrences).  Device spectra use the packed [lm][field] layout
(m >= 0 only; a real field's negative orders follow from conjugate
symmetry).  The reference's (L, 2L+1) [l, l+m] arrays are accepted and
returned by the API functions.
"""
import numpy as np
import torch

from .. import _native as N
from . import _backend as B
from . import _meshgen as G
from .spherical_spline import spherical_spline


def nlm(L):
    return L * (L + 1) // 2


def lm_index(L, l, m):
    return m * L - m * (m - 1) // 2 + (l - m)


def _lm_tables(L):
    ls, ms = [], []
    for m in range(L):
        for l in range(m, L):
            ls.append(l)
            ms.append(m)
    return np.array(ls), np.array(ms)


# arithmetic of the Legendre synthesis: "fp64" (reference precision),
# "fp32" (FP32 matrix cores, FP64 recurrence, FP64-compensated partial sums),
# "bf16xS" (S fixed-point bf16 digits per operand, FP32 accumulators) or
# "i8xS" (S int8 digits per operand, the Legendre digits tabulated once per
# plan, exact int32 accumulators)
_LEGENDRE = {"fp64": N.LEGENDRE_FP64, "fp32": N.LEGENDRE_FP32, "bf16x3": N.LEGENDRE_BF16X3,
             "bf16x4": N.LEGENDRE_BF16X4, "bf16x5": N.LEGENDRE_BF16X5, "bf16x6": N.LEGENDRE_BF16X6,
             "bf16x7": N.LEGENDRE_BF16X7, "i8x4": N.LEGENDRE_I8X4, "i8x5": N.LEGENDRE_I8X5,
             "i8x6": N.LEGENDRE_I8X6, "i8x7": N.LEGENDRE_I8X7, "i8x8": N.LEGENDRE_I8X8}


class SHT(object):
    """MWSS transform plan for bandlimit L (0 <= l < L)."""

    _cache = {}

    def __init__(self, L, legendre="fp64"):
        self.L = L
        self.legendre = legendre
        th = G.mwss_thetas(L)
        self.cos_t, self.sin_t = np.cos(th), np.sin(th)
        h = N.vp()
        import ctypes
        N.call("cmm_sht_plan_create", int(L), N.dptr(self.cos_t), N.dptr(self.sin_t), ctypes.byref(h))
        self._h = h.value
        if legendre not in _LEGENDRE:
            raise N.CmmError(f"legendre must be one of {sorted(_LEGENDRE)}")
        N.call("cmm_sht_set_legendre", self._h, _LEGENDRE[legendre])
        self._work = {}
        ls, ms = _lm_tables(L)
        self.l_of, self.m_of = ls, ms

    @classmethod
    def get(cls, L, legendre="fp64"):
        key = (L, torch.cuda.current_device(), legendre)
        p = cls._cache.get(key)
        if p is None:
            p = cls(L, legendre)
            cls._cache[key] = p
        return p

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            try:
                N.lib().cmm_sht_plan_destroy(h)
            except Exception:
                pass
            self._h = None

    def gl_nodes(self):
        x = np.zeros(self.L)
        w = np.zeros(self.L)
        N.call("cmm_sht_plan_gl", self._h, N.dptr(x), N.dptr(w))
        return x, w

    def work(self, nf):
        """Device workspace for nf fields on the current stream (transforms
        on different streams share the plan tables, never the workspace).
        A buffer first needed while the stream is being captured into a HIP
        graph belongs to that graph's memory pool and is not cached."""
        key = (nf, B.stream())
        t = self._work.get(key)
        if t is None:
            b = np.zeros(1, dtype=np.int64)
            N.call("cmm_sht_workspace_bytes", self._h, int(nf), N.lptr(b))
            t = B.empty(int(b[0]) // 8 + 1)
            if not torch.cuda.is_current_stream_capturing():
                self._work[key] = t
        return t

    def inverse(self, flm, nf=None, out=None):
        """flm (nlm, nf, 2) float64 (complex interleaved) -> (nf, L+1, 2L)."""
        if nf is None:
            nf = flm.shape[1]
        if out is None:
            out = B.empty(nf, self.L + 1, 2 * self.L)
        N.call("cmm_sht_inverse", self._h, flm.data_ptr(), int(nf), out.data_ptr(),
               self.work(nf).data_ptr(), B.stream())
        return out

    def forward(self, f, out=None):
        """f (nf, L+1, 2L) -> (nlm, nf, 2)."""
        f = f.reshape(-1, self.L + 1, 2 * self.L)
        nf = f.shape[0]
        if out is None:
            out = B.empty(nlm(self.L), nf, 2)
        N.call("cmm_sht_forward", self._h, f.data_ptr(), int(nf), out.data_ptr(),
               self.work(nf).data_ptr(), B.stream())
        return out

    def stream_operators(self, w_lm, apply_dinv=True, out=None):
        """omega_lm (nlm,1,2) -> (nlm, 12, 2): (r x grad) psi and (r x grad) of
        each component, psi = D_inv omega (spherical_harmonic_tools.py:14-19, 49-62)."""
        if out is None:
            out = B.empty(nlm(self.L), 12, 2)
        N.call("cmm_sht_stream_operators", self._h, w_lm.data_ptr(), int(bool(apply_dinv)),
               out.data_ptr(), B.stream())
        return out

    def scalar_operators(self, f_lm, out=None):
        if out is None:
            out = B.empty(nlm(self.L), 4, 2)
        N.call("cmm_sht_scalar_operators", self._h, f_lm.data_ptr(), out.data_ptr(), B.stream())
        return out

    # layout conversion (set-up / API only)
    def full_to_packed(self, c):
        """(L, 2L+1) [l, l+m] complex -> (nlm, 1, 2) device (reads m >= 0)."""
        ct = c if isinstance(c, torch.Tensor) else torch.from_numpy(np.asarray(c, dtype=np.complex128))
        ct = ct.to(B.device())
        l = torch.from_numpy(self.l_of).to(ct.device)
        m = torch.from_numpy(self.m_of).to(ct.device)
        v = ct[l, l + m]
        return torch.view_as_real(v.contiguous()).reshape(-1, 1, 2).contiguous()

    def packed_to_full(self, p, field=0):
        """(nlm, nf, 2) -> (L, 2L+1) complex with negative orders by symmetry."""
        L = self.L
        v = torch.view_as_complex(p[:, field, :].contiguous())
        out = torch.zeros(L, 2 * L + 1, dtype=torch.complex128, device=p.device)
        l = torch.from_numpy(self.l_of).to(p.device)
        m = torch.from_numpy(self.m_of).to(p.device)
        out[l, l + m] = v
        neg = m > 0
        sg = torch.where((m[neg] % 2) == 1, -1.0, 1.0).to(torch.float64)
        out[l[neg], l[neg] - m[neg]] = sg * torch.conj(v[neg])
        return out


# ----------------------------------------------------------------------------
# reference API
# ----------------------------------------------------------------------------
def operator_tables(L):
    """spherical_simulations.py:195-205: D_inv, L_plus, L_minus, L_z."""
    els = np.arange(0, L)
    els[0] = 1
    D_inv = -1 / (els * (els + 1))
    D_inv[0] = 0.
    with np.errstate(invalid="ignore"):
        L_plus = np.fromfunction(lambda l, m: np.sqrt(l * (l + 1) - (m - l) * (m - l - 1)), (L, 2 * L + 1))
        L_minus = np.fromfunction(lambda l, m: np.sqrt(l * (l + 1) - (m - l) * (m - l + 1)), (L, 2 * L + 1))
    L_plus[np.isnan(L_plus)] = 0.
    L_minus[np.isnan(L_minus)] = 0.
    L_z = np.fromfunction(lambda l, m: m - l, (L, 2 * L + 1))
    return D_inv, L_plus, L_minus, L_z


def MW_sampling(L, method="MWSS"):
    """spherical_harmonic_tools.py:21-28 (MWSS)."""
    if method != "MWSS":
        raise ValueError("only MWSS sampling is provided")
    return list(np.meshgrid(G.mwss_phis(L), G.mwss_thetas(L)))


def coeff_array(alms, L):
    """spherical_harmonic_tools.py:30-42 (ssht index l^2 + l + m)."""
    outs = np.zeros([L, 2 * L + 1], dtype="complex128")
    for l in range(L):
        outs[l, :2 * l + 1] = np.asarray(alms)[l * l:l * l + 2 * l + 1]
    return outs


def lm_array(coeffs, L):
    """spherical_harmonic_tools.py:44-47."""
    c = np.asarray(coeffs)
    return np.concatenate([c[l, :2 * l + 1] for l in range(L)])


def inv_Laplacian(samples, L, L_inv):
    """spherical_harmonic_tools.py:14-19: forward MWSS transform x D_inv ->
    (L, 2L+1) complex."""
    plan = SHT.get(L)
    f = B.to_dev(samples).reshape(1, L + 1, 2 * L)
    flm = plan.forward(f)
    full = plan.packed_to_full(flm) * B.to_dev(L_inv).reshape(-1, 1)
    return B.like(full, samples)


def angular_momentum(alms, L_plus, L_minus, L_z, temp):
    """spherical_harmonic_tools.py:49-62 on (L, 2L+1) arrays (API helper)."""
    xp = torch if isinstance(alms, torch.Tensor) else np
    temp *= 0
    temp[:, 1:] = alms[:, :-1]
    out_plus = L_plus * temp
    temp *= 0
    temp[:, :-1] = alms[:, 1:]
    out_minus = L_minus * temp
    out_x = (out_plus + out_minus) / 2
    out_y = 1j * (out_minus - out_plus) / 2
    return [1j * out_x, 1j * out_y, 1j * (L_z * alms)]


def proj(a, b):
    scl = (a[0] * b[0] + a[1] * b[1] + a[2] * b[2]) / np.sqrt(b[0] ** 2 + b[1] ** 2 + b[2] ** 2)
    return [scl * b[0], scl * b[1], scl * b[2]]


def tan_proj(U, R):
    pU = proj(U, R)
    return [U[0] - pU[0], U[1] - pU[1], U[2] - pU[2]]


def project_stream_lm(plan, w_lm, U, n, apply_dinv=True, fields=None, ops=None):
    """Device core of project_stream: psi (or omega with apply_dinv) packed
    (nlm,1,2) -> U level n jets + coefficients."""
    if U.stepping:
        U.shift()
    ops = plan.stream_operators(w_lm, apply_dinv, out=ops)
    fields = plan.inverse(ops, 12, out=fields)
    N.call("cmm_mwss_velocity_jets", U.mesh._h, fields.data_ptr(), U.level_vals(n).data_ptr(), B.stream())
    U.init_coeffs(n)
    return fields


def project_stream(alms, L, L_plus, L_minus, L_z, U, n, temp):
    """spherical_harmonic_tools.py:66-105: psi_lm (L, 2L+1) -> velocity
    u = r x grad psi and grad u on the MWSS mesh vertices -> U level n."""
    plan = SHT.get(L)
    project_stream_lm(plan, plan.full_to_packed(alms), U, n, apply_dinv=False)


def project_onto_S12(alms, L, mesh):
    """spherical_harmonic_tools.py:108-142: (L, 2L+1) coefficients -> scalar
    PS spline on the MWSS mesh ``mesh``."""
    plan = SHT.get(L)
    return project_onto_S12_lm(plan, plan.full_to_packed(alms), mesh)


def project_onto_S12_lm(plan, f_lm, mesh):
    ops = plan.scalar_operators(f_lm)
    fields = plan.inverse(ops, 4)
    vals = B.empty(4, mesh.nv)
    N.call("cmm_mwss_scalar_jets", mesh._h, fields.data_ptr(), vals.data_ptr(), B.stream())
    return spherical_spline(mesh, vals)


def enstrophy_energy(f, L, plan=None):
    """Enstrophy sum_lm |f_lm|^2 and energy 0.5 sum_l E_l of a real field f
    on the MWSS grid (L+1, 2L) (or a batch (nf, L+1, 2L)), over all orders
    (m < 0 by conjugate symmetry), with E_0 = |f_00|^2 and E_l = sum_m |f_lm|^2 /
    (l (l+1)): the diagnostics of figures/data_gen_convergence.py:58-61, 91-95
    (pysh.forward with Reality=False, energy_spectrum of
    spherical_harmonic_tools.py:157-163), from the device forward transform.
    Returns two floats (or two arrays for a batch)."""
    plan = SHT.get(L) if plan is None else plan
    fd = B.to_dev(f)
    single = fd.dim() == 2
    a = plan.forward(fd.reshape(-1, L + 1, 2 * L))  # (nlm, nf, 2)
    key = ("ens", plan._h)
    w = _diag_weights.get(key)
    if w is None:
        l = plan.l_of.astype(np.float64)
        mult = np.where(plan.m_of == 0, 1.0, 2.0)
        inv = np.where(l > 0, 1.0 / np.maximum(l * (l + 1), 1.0), 1.0)
        w = (torch.from_numpy(mult).to(a.device), torch.from_numpy(mult * inv).to(a.device))
        _diag_weights[key] = w
    p2 = (a * a).sum(-1)  # (nlm, nf)
    ens = (w[0][:, None] * p2).sum(0)
    en = 0.5 * (w[1][:, None] * p2).sum(0)
    ens, en = ens.cpu().numpy(), en.cpu().numpy()
    return (float(ens[0]), float(en[0])) if single else (ens, en)


_diag_weights = {}


def energy_spectrum(alms, L):
    """spherical_harmonic_tools.py:157-163."""
    a = np.asarray(alms.detach().cpu() if isinstance(alms, torch.Tensor) else alms)
    outs = [np.sum(a[0, :] * a[0, :].conjugate())]
    for ell in range(1, L):
        outs.append(np.sum(a[ell, :] * a[ell, :].conjugate()) / (ell * (ell + 1)))
    return np.array(outs)
