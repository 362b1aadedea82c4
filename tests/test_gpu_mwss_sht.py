"""GPU parity: MWSS structured-mesh location and velocity splines, the MWSS
spherical-harmonic transforms, spectral operators and project_stream."""
import numpy as np
import pytest
import torch

from helpers import random_unit, rel_err

pytestmark = pytest.mark.gpu

T_POS = 1e-11


def _grid_and_ring_points(L, seed):
    from oracle import meshes as om
    th, ph = om.mwss_thetas(L), om.mwss_phis(L)
    rng = np.random.default_rng(seed)
    # points on and next to rings / meridians exercise the 0.03 / 0.85 bands
    t = np.concatenate([th[rng.integers(0, L + 1, 3000)] + rng.uniform(-0.1, 1.1, 3000) * (np.pi / L),
                        th[rng.integers(0, L + 1, 2000)]])
    p = np.concatenate([ph[rng.integers(0, 2 * L, 3000)] + rng.uniform(0, 1, 3000) * (np.pi / L),
                        ph[rng.integers(0, 2 * L, 2000)]])
    t = np.clip(t, 0, np.pi)
    q = np.stack([np.sin(t) * np.cos(p), np.sin(t) * np.sin(p), np.cos(t)], 1)
    return np.concatenate([q, random_unit(20000, seed + 1)])


@pytest.mark.parametrize("L", [8, 16, 24, 64])
def test_mwss_locate(L):
    from cmm_s2.core import mesh_functions as mf
    from oracle import meshes as om
    pm = mf.structure_spherical_triangulation(L)
    o = om.StructuredTriangulation(L)
    assert np.array_equal(pm.simplices, o.simplices)
    assert np.array_equal(pm.msimplices, o.msimplices)
    assert np.array_equal(pm.vertices, o.vertices)
    q = _grid_and_ring_points(L, L)
    bo, to, vo = o.query(q)
    t, b = pm.locate(torch.from_numpy(q).cuda(), "N3")
    t = t.cpu().numpy().astype(np.int64)
    b = b.cpu().numpy()
    same = t == to
    # a point whose oracle barycentrics are all clear of zero has a unique
    # triangle: identical index required.  Ties (points on an edge or a grid
    # vertex, e.g. the exact grid points above) may pick any triangle that
    # contains the point.
    clear = bo.min(axis=1) > 1e-12
    assert same[clear].all()
    bad = ~same
    if bad.any():
        vs = pm.vertices[pm.simplices[t[bad]]]
        bg = np.linalg.solve(np.transpose(vs, (0, 2, 1)), q[bad][:, :, None])[:, :, 0]
        assert (bg.min(axis=1) > -1e-12).all()
    # Cramer barycentrics: det / det on triangles of area ~ (pi/L)^2 lose
    # log10(1/area) digits; FMA vs numpy rounding differs at that level
    assert np.abs(b[same] - bo[same]).max() < 1e-12


def test_split_locate_from_known_triangle():
    """The RK-feet adjoint rebuilds a stage point's triangle in closed form
    from the triangle the forward recomputation found (no angle / cell
    search).  Through the C ABI, on stencil quads (eps = 1e-5) of seeded
    points next to rings and meridians and in both polar caps, at L = 32:
    the split location from the known triangle gives the full search's
    triangle and sub-triangle, and its sub-triangle barycentrics and dual
    basis to rounding (the two inlined contexts may fuse the Cramer products
    differently)."""
    from cmm_s2 import _native as N
    from cmm_s2.core import _backend as B, mesh_functions as mf
    L = 32
    pm = mf.structure_spherical_triangulation(L)
    rng = np.random.default_rng(32)
    base = _grid_and_ring_points(L, 7)[::4]
    caps = np.concatenate([rng.uniform(0, np.pi / L, 600), np.pi - rng.uniform(0, np.pi / L, 600)])
    ph = rng.uniform(0, 2 * np.pi, caps.size)
    base = np.concatenate([base, np.stack([np.sin(caps) * np.cos(ph), np.sin(caps) * np.sin(ph), np.cos(caps)], 1)])
    a1 = np.cross(base, [0.3, -0.5, 0.8])
    a1 /= np.linalg.norm(a1, axis=1)[:, None]
    a2 = np.cross(base, a1)
    eps = 1e-5
    q = np.concatenate([base + sa * eps * a1 + sb * eps * a2 for sa in (-1, 1) for sb in (-1, 1)])
    q /= np.linalg.norm(q, axis=1)[:, None]
    n = q.shape[0]
    qd = torch.from_numpy(q).cuda()

    def run(known):
        tri = torch.empty(n, dtype=torch.int32, device="cuda")
        sub = torch.empty(n, dtype=torch.int32, device="cuda")
        cb, w = B.empty(n, 3), B.empty(n, 9)
        N.call("cmm_mwss_split_locate", pm._h, qd.data_ptr(), n, 3, 1, None if known is None else known.data_ptr(),
               tri.data_ptr(), sub.data_ptr(), cb.data_ptr(), w.data_ptr(), B.stream())
        return tri, sub, cb, w

    t0, s0, c0, w0 = run(None)
    t1, s1, c1, w1 = run(t0)
    tl, _ = pm.locate(qd, "N3")
    assert torch.equal(t0, tl.to(torch.int32)) and int((t0 < 0).sum()) == 0
    tri = t0.cpu().numpy()
    Tu = 4 * L * (L - 1)
    assert (tri < 2 * L).sum() > 100 and (tri >= Tu - 2 * L).sum() > 100  # both polar caps
    assert ((tri >= 2 * L) & (tri < Tu - 2 * L)).sum() > 1000
    assert len(set(s0.cpu().numpy().tolist())) == 6
    assert torch.equal(t0, t1)
    # the sub-triangle is (argmin, argmax) of the barycentrics: it may only
    # differ where two of them tie to rounding
    c0h, c1h = c0.cpu().numpy(), c1.cpu().numpy()
    diff = (s0 != s1).cpu().numpy()
    if diff.any():
        b, _, _ = pm.query(q[diff])
        bs = np.sort(b, axis=1)
        assert (np.minimum(bs[:, 1] - bs[:, 0], bs[:, 2] - bs[:, 1]) < 1e-13).all()
    same = ~diff
    assert diff.sum() <= 8
    assert np.abs(c0h[same] - c1h[same]).max() < 1e-13
    w0h, w1h = w0.cpu().numpy(), w1.cpu().numpy()
    assert np.abs(w0h - w1h).max() < 1e-13 * np.abs(w0h).max()


def test_velocity_spline_eval():
    from cmm_s2.core import mesh_functions as mf
    from cmm_s2.core.spherical_spline import spline_interp_velocity
    from oracle import meshes as om, splines as osp
    L = 16
    pm = mf.structure_spherical_triangulation(L)
    o = om.StructuredTriangulation(L)
    rng = np.random.default_rng(0)
    vals = rng.standard_normal((3, 3, 4, pm.nv))
    U = spline_interp_velocity(pm, torch.from_numpy(vals).cuda(), None, ts=[0.0, 0.1, 0.2])
    Uo = osp.VelocitySpline(o, vals.copy(), np.zeros((3, 3, 19, pm.ntri)), ts=np.array([0.0, 0.1, 0.2]))
    for i in range(3):
        U.init_coeffs(i)
        Uo.init_coeffs(i)
    assert rel_err(U.coeffs.cpu().numpy(), Uo.coeffs) < T_POS
    q = random_unit(20000, 1).T.copy()
    for nVs, t in [(1, 0.05), (2, 0.17), (3, 0.3), (3, 0.25)]:
        U.nVs = Uo.nVs = nVs
        got = U(t, 0.1, torch.from_numpy(q).cuda()).cpu().numpy()
        ref = Uo(t, 0.1, q)
        assert rel_err(got, ref) < T_POS


@pytest.mark.parametrize("L", [16, 32, 48])
def test_sht_inverse_forward(L):
    from cmm_s2.core.spherical_harmonic_tools import SHT, nlm
    from oracle import sht as osh
    plan = SHT.get(L)
    x, w = plan.gl_nodes()
    po = osh.plan(L)
    assert np.abs(x - po.xgl).max() < 1e-15 and rel_err(w, po.wgl) < 1e-14
    rng = np.random.default_rng(L)
    nf = 3
    a = np.zeros((nf, L, L), complex)
    for f in range(nf):
        for l in range(L):
            a[f, l, :l + 1] = rng.standard_normal(l + 1) + 1j * rng.standard_normal(l + 1)
        a[f, :, 0] = a[f, :, 0].real
    packed = np.zeros((nlm(L), nf, 2))
    for m in range(L):
        for l in range(m, L):
            i = m * L - m * (m - 1) // 2 + (l - m)
            packed[i, :, 0] = a[:, l, m].real
            packed[i, :, 1] = a[:, l, m].imag
    g = plan.inverse(torch.from_numpy(packed).cuda()).cpu().numpy()
    ref = osh.inverse_lm(a, L)
    assert rel_err(g, ref) < 1e-12
    back = plan.forward(torch.from_numpy(ref).cuda()).cpu().numpy()
    assert rel_err(back, packed) < 1e-12
    # forward on non-band-limited data follows the same (documented) algorithm
    f = rng.standard_normal((2, L + 1, 2 * L))
    f[:, 0, :] = f[:, 0, :1]
    f[:, -1, :] = f[:, -1, :1]
    fo = osh.forward_lm(f, L)
    fg = plan.forward(torch.from_numpy(f).cuda()).cpu().numpy()
    for fi in range(2):
        ref_p = np.zeros((nlm(L), 2))
        for m in range(L):
            for l in range(m, L):
                i = m * L - m * (m - 1) // 2 + (l - m)
                ref_p[i] = [fo[fi, l, m].real, fo[fi, l, m].imag]
        assert rel_err(fg[:, fi], ref_p) < 1e-12


@pytest.mark.parametrize("L", [16, 24, 48])
def test_sht_forward_non_band_limited(L):
    """The device forward transform on input that is not band-limited: pure
    theta-Nyquist rings cos(L theta) cos(m phi), odd orders sin(k theta)
    cos(m phi) and steep smooth fields, against the exact projection of
    their split-Nyquist trigonometric interpolant computed independently
    (numpy FFTs, 4L-node Gauss-Legendre quadrature, scipy sph_harm_y)."""
    from cmm_s2.core.spherical_harmonic_tools import SHT, nlm
    from helpers import interpolant_projection, mwss_samples
    plan = SHT.get(L)
    funs = [lambda T, P: np.cos(L * T) * np.cos(2 * P), lambda T, P: np.cos(L * T),
            lambda T, P: np.sin((L - 1) * T) * np.cos(3 * P),
            lambda T, P: np.tanh(8 * (np.cos(T) - 0.3)) * (1 + 0.2 * np.sin(T) * np.cos(3 * P)),
            lambda T, P: np.exp(-((np.sin(T) * np.cos(P) - 1) ** 2 + (np.sin(T) * np.sin(P)) ** 2
                                  + np.cos(T) ** 2) * 16)]
    f = np.stack([mwss_samples(fn, L) for fn in funs])
    got = plan.forward(torch.from_numpy(f).cuda()).cpu().numpy()
    for fi in range(len(funs)):
        ref = interpolant_projection(f[fi], L)
        ref_p = np.zeros((nlm(L), 2))
        for m in range(L):
            for l in range(m, L):
                i = m * L - m * (m - 1) // 2 + (l - m)
                ref_p[i] = [ref[l, m].real, ref[l, m].imag]
        assert np.abs(got[:, fi] - ref_p).max() < 1e-12 * np.abs(ref_p).max(), fi


@pytest.mark.parametrize("L", [27, 37, 43, 384, 768, 1000, 1024, 2048])
def test_sht_any_bandlimit(L):
    """Bandlimits whose ring length 2L has prime factors other than 2 and 3
    (2 x 37; 2 x 43, whose L + 1 rings fill whole steps of 4 with no zero
    padding ring; 2^4 x 5^3, the reference's upsampling grids Lf = 1000), the
    largest plan (L = 2048, the reference's Poisson test up to 2048), and the
    ring FFTs that run in place on one LDS buffer with radix-2 and radix-3
    passes and up to 4 butterflies per thread (2L = 54 = 2 x 3^3, starting
    with a radix-2 pass; 768 = 4^4 x 3; 1536 = 4^4 x 2 x 3; 2048 = 4^5 x 2):
    synthesis of single Y_lm against scipy's sph_harm_y on the grid, the
    Rossby-Haurwitz (5,4) Poisson closed form of the reference's test
    (tests/test_spherical_harmonic_tools.py:48-56), and analysis of random
    band-limited spectra returning them."""
    from scipy.special import sph_harm_y
    from cmm_s2.core.spherical_harmonic_tools import SHT, nlm
    from helpers import mwss_samples
    plan = SHT.get(L)
    th = np.pi * np.arange(L + 1) / L
    ph = np.pi * np.arange(2 * L) / L
    # scipy's sph_harm_y overflows to NaN for low orders at degrees beyond
    # ~900; those picks stay below it
    picks = [(0, 0), (5, 4), (min(L - 1, 600), 1), (L - 1, L - 1), (min(L - 1, 503), min(L - 2, 333))]
    packed = np.zeros((nlm(L), len(picks), 2))
    for f, (l, m) in enumerate(picks):
        packed[m * L - m * (m - 1) // 2 + (l - m), f] = [1.0, 0.5]
    g = plan.inverse(torch.from_numpy(packed).cuda()).cpu().numpy()
    rows = np.arange(0, L + 1, max(1, L // 64))
    for f, (l, m) in enumerate(picks):
        Y = sph_harm_y(l, m, th[rows][:, None], ph[None, :])
        assert np.isfinite(Y).all(), (l, m)
        ref = ((1.0 + 0.5j) * Y).real * (1 if m == 0 else 2)
        assert np.abs(g[f, rows] - ref).max() < 1e-11 * max(1.0, np.abs(ref).max()), (l, m)
    # Delta psi = w for the RH (5,4) wave: psi = -z((1-z^2)^2 - 8(1-z^2)x^2 + 8x^4)
    w = mwss_samples(lambda T, P: 30 * np.cos(T) * np.cos(4 * P) * np.sin(T) ** 4, L)
    w_lm = plan.forward(torch.from_numpy(w).cuda().reshape(1, L + 1, 2 * L))
    dinv = np.zeros(L)
    dinv[1:] = -1.0 / (np.arange(1, L) * (np.arange(1, L) + 1.0))
    psi = plan.inverse(w_lm * torch.from_numpy(dinv[plan.l_of]).cuda()[:, None, None])[0].cpu().numpy()
    x, y, z = [mwss_samples(f, L) for f in (lambda T, P: np.sin(T) * np.cos(P), lambda T, P: np.sin(T) * np.sin(P),
                                               lambda T, P: np.cos(T) + 0 * P)]
    sol = -z * ((1 - z ** 2) ** 2 - 8 * (1 - z ** 2) * x ** 2 + 8 * x ** 4)
    assert rel_err(psi, sol) < 1e-11
    rng = np.random.default_rng(L)
    a = rng.standard_normal((nlm(L), 1, 2)) * (1.0 + plan.l_of[:, None, None]) ** -1.0
    a[plan.m_of == 0, :, 1] = 0.0
    back = plan.forward(plan.inverse(torch.from_numpy(a).cuda())).cpu().numpy()
    assert rel_err(back, a) < 1e-11


def test_project_stream_matches_oracle():
    from cmm_s2.core import mesh_functions as mf
    from cmm_s2.core.spherical_spline import spline_interp_velocity
    from cmm_s2.core import spherical_harmonic_tools as st
    from oracle import meshes as om, splines as osp, sht as osh, fields as ofl
    L = 32
    Phi, The = np.meshgrid(om.mwss_phis(L), om.mwss_thetas(L))
    w0 = ofl.rossby_wave(Phi, The) + 0.3 * ofl.gaussian_vortex(Phi, The)
    D_inv, Lp, Lm, Lz = osh.operator_tables(L)
    psi_o = osh.inv_Laplacian(w0, L, D_inv)
    o = om.StructuredTriangulation(L)
    Uo = osp.VelocitySpline(o, np.zeros((3, 3, 4, o.vertices.shape[0])),
                            np.zeros((3, 3, 19, len(o.simplices))), ts=np.zeros(3))
    osh.project_stream(psi_o, L, Lp, Lm, Lz, Uo, 0, 0 * psi_o)
    pm = mf.structure_spherical_triangulation(L)
    U = spline_interp_velocity(pm, torch.zeros(3, 3, 4, pm.nv, dtype=torch.float64, device="cuda"),
                               None, ts=[0, 0, 0])
    plan = st.SHT.get(L)
    w_lm = plan.forward(torch.from_numpy(w0).cuda().reshape(1, L + 1, 2 * L))
    st.project_stream_lm(plan, w_lm, U, 0, apply_dinv=True)
    gv = U.level_vals(0).cpu().numpy()
    assert rel_err(gv, Uo.vals[0]) < 1e-11
    # reference-API path from (L, 2L+1) coefficients
    psi_g = st.inv_Laplacian(torch.from_numpy(w0).cuda(), L, D_inv).cpu().numpy()
    assert rel_err(psi_g, psi_o) < 1e-12
    U2 = spline_interp_velocity(pm, torch.zeros(3, 3, 4, pm.nv, dtype=torch.float64, device="cuda"),
                                None, ts=[0, 0, 0])
    st.project_stream(psi_o, L, Lp, Lm, Lz, U2, 0, None)
    assert rel_err(U2.level_vals(0).cpu().numpy(), Uo.vals[0]) < 1e-11


def test_project_onto_S12_and_sampling():
    from cmm_s2.core import mesh_functions as mf
    from cmm_s2.core import spherical_harmonic_tools as st, dynamical_fields as df
    from cmm_s2.core.spherical_simulations import VorticitySampler
    from oracle import meshes as om, sht as osh, fields as ofl
    L = 24
    a = ofl.random_sh_coefficients(2, 8, L)
    assert np.array_equal(a, df.random_sh_coefficients(2, 8, L))
    full = osh.lm_to_full(a, L)
    o = om.StructuredTriangulation(L)
    so = osh.project_onto_S12(full, L, o)
    pm = mf.structure_spherical_triangulation(L)
    sg = st.project_onto_S12(full, L, pm)
    assert rel_err(sg.vals.cpu().numpy(), so.vals) < 1e-11
    # sampling w0(X) + 2 rot (cos th(X) - cos th) at arbitrary points
    Lg = 12
    g = mf.structure_spherical_triangulation(Lg)
    X = random_unit((Lg + 1) * 2 * Lg, 3)
    samp = VorticitySampler(sg, g, 2 * np.pi)
    got = samp(torch.from_numpy(X).cuda()).cpu().numpy().reshape(Lg + 1, 2 * Lg)
    ph, th = (np.arctan2(X[:, 1], X[:, 0]) + 2 * np.pi) % (2 * np.pi), np.arctan2(np.hypot(X[:, 0], X[:, 1]), X[:, 2])
    ref = so(ph.reshape(Lg + 1, 2 * Lg), th.reshape(Lg + 1, 2 * Lg))
    ref = ref + 2 * 2 * np.pi * np.cos(th.reshape(Lg + 1, -1)) - 2 * 2 * np.pi * np.cos(om.mwss_thetas(Lg))[:, None]
    assert rel_err(got, ref) < 1e-11
    # analytic vorticity kernels
    for name in ["rossby_wave", "zonal_jet", "gaussian_vortex", "perturbed_zonal_jet", "multi_jet",
                 "perturbed_rossby_wave"]:
        samp = VorticitySampler(getattr(df, name), g, 0.0)
        got = samp(torch.from_numpy(X).cuda()).cpu().numpy()
        ref = getattr(ofl, name)(ph, th)
        assert rel_err(got, ref) < 1e-12, name


@pytest.mark.parametrize("L", [64, 256])
def test_fp32_legendre_synthesis_error(L):
    """Mixed-precision synthesis (FP32 matrix cores, FP64 recurrence,
    FP64-compensated 32-degree partial sums) against the FP64 synthesis on
    the same spectrum: relative max-norm error per field below 2e-6, and
    clearly not the FP64 path (error above 1e-12)."""
    from cmm_s2.core import spherical_harmonic_tools as st
    p64, p32 = st.SHT.get(L, "fp64"), st.SHT.get(L, "fp32")
    g = torch.Generator(device="cuda").manual_seed(L)
    ls = torch.from_numpy(p64.l_of.astype(np.float64)).cuda()
    # l^-1.5 amplitude decay, 12 fields as in the velocity projection
    flm = torch.randn(st.nlm(L), 12, 2, dtype=torch.float64, device="cuda", generator=g)
    flm *= (1.0 + ls)[:, None, None] ** -1.5
    f64 = p64.inverse(flm)
    f32 = p32.inverse(flm)
    for k in range(12):
        e = float((f32[k] - f64[k]).abs().max() / f64[k].abs().max())
        assert 1e-12 < e < 2e-6, (k, e)


_SPLIT_BOUND = {"bf16x3": 1e-4, "bf16x4": 1e-6, "bf16x5": 1e-9, "bf16x6": 1e-11, "bf16x7": 2e-13}


@pytest.mark.parametrize("L,nf", [(16, 1), (27, 5), (64, 12), (256, 12), (768, 3), (1024, 20)])
def test_split_bf16_legendre_synthesis_error(L, nf):
    """Split-BF16 synthesis (S bf16 fixed-point digits per operand on bf16
    matrix cores, exact FP32 sums of the four leading weights, FP64 flush
    every 192 degrees per parity) against the FP64 synthesis on the same
    spectrum: relative max-norm error per field below the bound of each S,
    falling with S.  L = 768 and 1024 pass the flush, nf = 20 runs two field
    groups."""
    from cmm_s2.core import spherical_harmonic_tools as st
    p64 = st.SHT.get(L, "fp64")
    g = torch.Generator(device="cuda").manual_seed(1000 + L)
    ls = torch.from_numpy(p64.l_of.astype(np.float64)).cuda()
    flm = torch.randn(st.nlm(L), nf, 2, dtype=torch.float64, device="cuda", generator=g)
    flm *= (1.0 + ls)[:, None, None] ** -1.5
    f64 = p64.inverse(flm)
    prev = None
    for mode, bound in _SPLIT_BOUND.items():
        fs = st.SHT.get(L, mode).inverse(flm)
        errs = [float((fs[k] - f64[k]).abs().max() / f64[k].abs().max()) for k in range(nf)]
        print(L, nf, mode, "max rel err", max(errs), flush=True)
        assert max(errs) < bound, (mode, max(errs))
        if prev is not None:
            assert max(errs) < prev, (mode, max(errs), prev)
        prev = max(errs)


_I8_BOUND = {"i8x4": 1e-6, "i8x5": 5e-9, "i8x6": 2e-11, "i8x7": 1e-13, "i8x8": 1e-13}


@pytest.mark.parametrize("L,nf", [(16, 1), (27, 5), (64, 12), (255, 3), (512, 12), (1024, 20)])
def test_int8_table_legendre_synthesis_error(L, nf):
    """Int8 digit-table synthesis (S balanced int8 digits per operand, the
    Legendre digits tabulated once per plan, exact int32 sums by weight on
    the int8 matrix cores) against the FP64 synthesis on the same spectrum:
    relative max-norm error per field below the bound of each S, falling
    with S down to the FP64 rounding level (S = 7, 8).  Odd L = 27 and 255
    exercise partial ring tiles and the equator-free ring sets, nf = 20 two
    field groups."""
    from cmm_s2.core import spherical_harmonic_tools as st
    p64 = st.SHT.get(L, "fp64")
    g = torch.Generator(device="cuda").manual_seed(2000 + L)
    ls = torch.from_numpy(p64.l_of.astype(np.float64)).cuda()
    flm = torch.randn(st.nlm(L), nf, 2, dtype=torch.float64, device="cuda", generator=g)
    flm *= (1.0 + ls)[:, None, None] ** -1.5
    f64 = p64.inverse(flm)
    prev = None
    for mode, bound in _I8_BOUND.items():
        fs = st.SHT.get(L, mode).inverse(flm)
        errs = [float((fs[k] - f64[k]).abs().max() / f64[k].abs().max()) for k in range(nf)]
        print(L, nf, mode, "max rel err", max(errs), flush=True)
        assert max(errs) < bound, (mode, max(errs))
        if prev is not None and mode != "i8x8":
            assert max(errs) < prev, (mode, max(errs), prev)
        prev = max(errs)


@pytest.mark.parametrize("L,nf", [(8, 2), (33, 3), (64, 12)])
def test_int8_table_synthesis_against_oracle(L, nf):
    """The int8 synthesis with 7 and 8 digits against the oracle's FP64
    synthesis (the FP64 tolerance), 5 digits within its bound; the pole
    rings (summed in FP64) included."""
    from cmm_s2.core import spherical_harmonic_tools as st
    from oracle import sht as osh
    rng = np.random.default_rng(L)
    p64 = st.SHT.get(L, "fp64")
    a = np.zeros((nf, L, L), complex)
    for f in range(nf):
        a[f] = rng.standard_normal((L, L)) + 1j * rng.standard_normal((L, L))
    a *= np.tril(np.ones((L, L)))[None]  # m <= l
    a[:, :, 0] = a[:, :, 0].real
    ref = osh.inverse_lm(a, L)
    flm = np.stack([a[:, p64.l_of, p64.m_of].real.T, a[:, p64.l_of, p64.m_of].imag.T], -1)
    for mode, bound in [("i8x5", 5e-9), ("i8x7", T_POS), ("i8x8", T_POS)]:
        got = st.SHT.get(L, mode).inverse(torch.from_numpy(np.ascontiguousarray(flm)).cuda()).cpu().numpy()
        for f in range(nf):
            assert rel_err(got[f], ref[f]) < bound, (mode, f, rel_err(got[f], ref[f]))


def test_enstrophy_energy_against_oracle():
    """Device enstrophy / energy (forward transform + weighted sums) against
    the oracle's restatement of figures/data_gen_convergence.py:58-95 on a
    random real field, and on a batch."""
    from oracle import sht as osht
    from cmm_s2.core import spherical_harmonic_tools as st
    rng = np.random.default_rng(8)
    L = 24
    fs = np.stack([rng.standard_normal((L + 1, 2 * L)) for _ in range(3)])
    for k in range(3):
        e, n = st.enstrophy_energy(fs[k], L)
        e0, n0 = osht.enstrophy(fs[k], L), osht.energy(fs[k], L)
        assert abs(e - e0) <= 1e-12 * e0 and abs(n - n0) <= 1e-12 * n0, (k, e, e0, n, n0)
    eb, nb = st.enstrophy_energy(fs, L)
    for k in range(3):
        assert abs(eb[k] - osht.enstrophy(fs[k], L)) <= 1e-12 * eb[k]


def test_conservation_errors_of_a_steady_run():
    """BarotropicCMM.conservation_errors on the steady zonal jet: at t = 0
    the absolute-vorticity enstrophy and energy equal the oracle's on the same
    grid, and after the run the relative errors are small (the exact solution
    is steady, so max_rel_err is the map error)."""
    from oracle import sht as osht, fields as ofl, meshes as om
    from cmm_s2.core import dynamical_fields as df, mesh_functions as mf
    from cmm_s2.core.spherical_simulations import BarotropicCMM
    ico = mf.icosahedral_mesh(4)
    mesh = mf.spherical_triangulation(ico.points, simplices=ico.simplices)
    L, Lo = 32, 48
    sim = BarotropicCMM(L, 12, 12 / L, mesh, df.zonal_jet, 2 * np.pi, n_maps=4)
    sim.run()
    th = om.mwss_thetas(Lo)
    ph = np.arange(2 * Lo) * np.pi / Lo
    PH, TH = np.meshgrid(ph, th)
    w_true = ofl.zonal_jet(PH, TH)
    r = sim.conservation_errors(Lo, true_vorticity=w_true)
    wabs = w_true + 2 * 2 * np.pi * np.cos(TH)
    assert abs(r["enstrophy_0"] - osht.enstrophy(wabs, Lo)) <= 1e-10 * r["enstrophy_0"]
    assert abs(r["energy_0"] - osht.energy(wabs, Lo)) <= 1e-10 * r["energy_0"]
    # measured: 4.9e-5, 8.7e-5 and 1.7e-2 (the jet's steep flanks on ico-4 / L = 32)
    assert r["enstrophy_rel_err"] < 1e-3 and r["energy_rel_err"] < 1e-3 and r["max_rel_err"] < 5e-2, r
