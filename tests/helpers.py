"""Shared test helpers (oracle <-> product index mapping, tolerances)."""
import numpy as np


def rel_err(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.max(np.abs(a - b)) / max(1e-300, np.max(np.abs(b))))


def tri_map(prod_simplices, oracle_simplices):
    """oracle row -> product row, by sorted vertex triples."""
    key = {tuple(sorted(t)): i for i, t in enumerate(np.asarray(prod_simplices).tolist())}
    return np.array([key[tuple(sorted(t))] for t in np.asarray(oracle_simplices).tolist()])


def random_unit(n, seed):
    g = np.random.default_rng(seed).standard_normal((n, 3))
    return g / np.linalg.norm(g, axis=1)[:, None]


def perturbed_identity_jets(vertices, seed, amp=0.02):
    """Smooth map jets: identity plus a low-degree polynomial perturbation,
    vals[c,0] = f_c(v), vals[d,1+c] = d_d f_c (tangential gradient)."""
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((3, 3, 3)) * amp
    A = A + A.transpose(0, 2, 1)
    V = vertices
    nv = len(V)
    jets = np.zeros((3, 4, nv))
    for c in range(3):
        f = V[:, c] + np.einsum("ni,ij,nj->n", V, A[c], V)
        g = np.eye(3)[c][None, :] + 2 * V @ A[c]
        g = g - np.einsum("ni,ni->n", g, V)[:, None] * V
        jets[c, 0] = f
        jets[:, 1 + c] = g.T
    return jets


def interpolant_projection(f, L, split_nyquist=True, nodes=None):
    """Spherical-harmonic coefficients [l, m >= 0] (L, L) complex of the
    continuous field that MWSS samples f (L+1, 2L) define, by exact
    projection: along each ring the trigonometric interpolant in phi (orders
    m < L); for each order the 2L-periodic trigonometric interpolant in
    theta of the ring coefficients, extended to the full circle with
    F_m(2 pi - theta) = (-1)^m F_m(theta), the Nyquist coefficient split
    evenly between e^{iL theta} and e^{-iL theta} (a cos(L theta) term; with
    split_nyquist False it is put on e^{iL theta} alone).  Computed
    independently of the transforms under test: numpy FFTs, Gauss-Legendre
    quadrature with `nodes` (4L) points in cos theta and scipy's
    sph_harm_y."""
    from scipy.special import sph_harm_y
    f = np.asarray(f, dtype=np.float64)
    n2 = 2 * L
    Fm = np.fft.rfft(f, axis=1)[:, :L] / n2                      # (L+1, L)
    m = np.arange(L)
    s = np.zeros((n2, L), dtype=np.complex128)
    s[:L + 1] = Fm
    s[L + 1:] = ((-1.0) ** m)[None, :] * Fm[1:L][::-1]           # s[2L - t] = (-1)^m F[t]
    c = np.fft.fft(s, axis=0) / n2                               # c[k]: coefficient of e^{i k theta}
    k = np.arange(n2)
    k[k > L] -= n2
    x, w = np.polynomial.legendre.leggauss(nodes or 4 * L)
    th = np.arccos(x)
    E = np.exp(1j * np.outer(th, k))                             # (nx, 2L)
    if split_nyquist:
        E[:, L] = np.cos(L * th)
    G = E @ c                                                    # (nx, L): G_m(theta_j)
    out = np.zeros((L, L), dtype=np.complex128)
    for mm in range(L):
        ls = np.arange(mm, L)
        P = sph_harm_y(ls[:, None], mm, th[None, :], 0.0).real   # (L - m, nx)
        out[mm:, mm] = 2 * np.pi * (P * w[None, :]) @ G[:, mm]
    return out


def theta_interpolant(f, L, theta, split_nyquist=True):
    """The theta interpolant of interpolant_projection evaluated at `theta`:
    (len(theta), L) complex, order m in the columns."""
    f = np.asarray(f, dtype=np.float64)
    n2 = 2 * L
    Fm = np.fft.rfft(f, axis=1)[:, :L] / n2
    m = np.arange(L)
    s = np.zeros((n2, L), dtype=np.complex128)
    s[:L + 1] = Fm
    s[L + 1:] = ((-1.0) ** m)[None, :] * Fm[1:L][::-1]
    c = np.fft.fft(s, axis=0) / n2
    k = np.arange(n2)
    k[k > L] -= n2
    E = np.exp(1j * np.outer(theta, k))
    if split_nyquist:
        E[:, L] = np.cos(L * np.asarray(theta))
    return E @ c, s


def mwss_samples(fun, L):
    """fun(theta, phi) on the MWSS grid (L+1, 2L), theta_t = pi t / L,
    phi_p = pi p / L."""
    th = np.pi * np.arange(L + 1) / L
    ph = np.pi * np.arange(2 * L) / L
    T, P = np.meshgrid(th, ph, indexing="ij")
    return fun(T, P)
