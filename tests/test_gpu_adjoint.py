"""GPU checks of the adjoint entry points and of the loop gradient.

Linear stages (SHT synthesis / analysis, spectral operators, grid -> jets,
assemblies): the dot-product identity <A x, y> = <x, A^T y> on random x, y,
to round-off.  Nonlinear stages (vorticity sampling, composition, stencil
projection, RK4 feet): the adjoint's directional derivative against a central
difference of the forward entry point along a random smooth direction.  The
loop gradient: central differences of J along random directions."""
import numpy as np
import pytest
import torch

from helpers import perturbed_identity_jets

pytestmark = pytest.mark.gpu


def _dot(a, b):
    return float((a.reshape(-1) * b.reshape(-1)).sum())


def _rand(*shape, seed=0):
    g = torch.Generator().manual_seed(seed)
    return torch.randn(*shape, generator=g, dtype=torch.float64).cuda()


def _close(a, b, tol):
    assert abs(a - b) <= tol * max(abs(a), abs(b), 1e-300), (a, b)


def _work(records, key_range=None):
    """(pointer, bytes) of a grouped scatter-add workspace for ``records``
    records with keys below ``key_range``, None = atomics."""
    import ctypes
    from cmm_s2 import _native as N
    if records is None:
        return None, 0
    b = ctypes.c_int64(0)
    N.call("cmm_scatter_workspace_bytes", int(records), int(key_range), ctypes.byref(b))
    t = torch.empty(b.value, dtype=torch.uint8, device="cuda")
    _work.keep = t
    return t.data_ptr(), b.value


@pytest.mark.parametrize("L", [16, 24, 64, 256, 512])
def test_sht_transposes(L):
    from cmm_s2 import _native as N
    from cmm_s2.core._backend import stream
    from cmm_s2.core.spherical_harmonic_tools import SHT, nlm
    p = SHT.get(L)
    for nf in (1, 4, 12):
        a = _rand(nlm(L), nf, 2, seed=L + nf)
        fb = _rand(nf, L + 1, 2 * L, seed=2 * L + nf)
        Sa = p.inverse(a, nf)
        ab = torch.empty_like(a)
        N.call("cmm_sht_inverse_adjoint", p._h, fb.data_ptr(), nf, ab.data_ptr(), p.work(nf).data_ptr(), stream())
        _close(_dot(Sa, fb), _dot(a, ab), 1e-12)
        f = _rand(nf, L + 1, 2 * L, seed=3 * L + nf)
        Af = p.forward(f)
        fbar = torch.empty_like(f)
        N.call("cmm_sht_forward_adjoint", p._h, a.data_ptr(), nf, fbar.data_ptr(), p.work(nf).data_ptr(), stream())
        _close(_dot(Af, a), _dot(f, fbar), 1e-12)


@pytest.mark.parametrize("L", [16, 48])
def test_spectral_operator_transposes(L):
    from cmm_s2 import _native as N
    from cmm_s2.core._backend import stream
    from cmm_s2.core.spherical_harmonic_tools import SHT, nlm
    p = SHT.get(L)
    w = _rand(nlm(L), 1, 2, seed=1)
    for dinv in (0, 1):
        ob = _rand(nlm(L), 12, 2, seed=2 + dinv)
        Ow = p.stream_operators(w, bool(dinv))
        wb = torch.empty_like(w)
        scratch = torch.empty(nlm(L), 3, 2, dtype=torch.float64, device="cuda")
        N.call("cmm_sht_stream_operators_adjoint", p._h, ob.data_ptr(), dinv, wb.data_ptr(), scratch.data_ptr(),
               stream())
        _close(_dot(Ow, ob), _dot(w, wb), 1e-12)
    ob4 = _rand(nlm(L), 4, 2, seed=7)
    Sw = p.scalar_operators(w)
    wb = torch.empty_like(w)
    N.call("cmm_sht_scalar_operators_adjoint", p._h, ob4.data_ptr(), wb.data_ptr(), stream())
    _close(_dot(Sw, ob4), _dot(w, wb), 1e-12)


def test_jets_and_assembly_transposes():
    from cmm_s2 import _native as N
    from cmm_s2.core._backend import stream
    from cmm_s2.core import mesh_functions as mf
    L = 16
    mu = mf.structure_spherical_triangulation(L)
    G = (L + 1) * 2 * L
    for nf, fwd, adj in ((12, "cmm_mwss_velocity_jets", "cmm_mwss_velocity_jets_adjoint"),
                         (4, "cmm_mwss_scalar_jets", "cmm_mwss_scalar_jets_adjoint")):
        F = _rand(nf, G, seed=nf)
        vb = _rand(nf, mu.nv, seed=nf + 1)
        v = torch.empty(nf, mu.nv, dtype=torch.float64, device="cuda")
        N.call(fwd, mu._h, F.data_ptr(), v.data_ptr(), stream())
        Fb = torch.empty_like(F)
        N.call(adj, mu._h, vb.data_ptr(), Fb.data_ptr(), stream())
        _close(_dot(v, vb), _dot(F, Fb), 1e-12)
    for nc in (1, 3):
        vals = _rand(4 * nc, mu.nv, seed=10 + nc)
        pb = _rand(mu.ntri, 36 * nc, seed=20 + nc)
        pk = torch.empty(mu.ntri, 36 * nc, dtype=torch.float64, device="cuda")
        N.call("cmm_mwss_assemble", mu._h, vals.data_ptr(), nc, None, pk.data_ptr(), stream())
        outs = []
        for gather in (False, True):  # atomics, then the two-phase gather
            vb = torch.zeros_like(vals)
            ws = torch.empty(96 * nc * mu.ntri, dtype=torch.uint8, device="cuda") if gather else None
            N.call("cmm_mwss_assemble_adjoint", mu._h, pb.data_ptr(), nc, vb.data_ptr(),
                   None if ws is None else ws.data_ptr(), 0 if ws is None else ws.numel(), stream())
            _close(_dot(pk, pb), _dot(vals, vb), 1e-12)
            outs.append(vb)
        assert torch.allclose(outs[0], outs[1], rtol=1e-12, atol=1e-12)
    ico = mf.icosahedral_mesh(3)
    mesh = mf.spherical_triangulation(ico.points, simplices=ico.simplices)
    vals = _rand(12, mesh.nv, seed=31)
    pb = _rand(mesh.ntri, 108, seed=32)
    pk = torch.empty(mesh.ntri, 108, dtype=torch.float64, device="cuda")
    N.call("cmm_tri_assemble", mesh._h, vals.data_ptr(), 3, None, pk.data_ptr(), stream())
    outs = []
    for gather in (False, True):
        vb = torch.zeros_like(vals)
        ws = torch.empty(288 * mesh.ntri, dtype=torch.uint8, device="cuda") if gather else None
        N.call("cmm_tri_assemble_adjoint", mesh._h, pb.data_ptr(), vb.data_ptr(),
               None if ws is None else ws.data_ptr(), 0 if ws is None else ws.numel(), stream())
        _close(_dot(pk, pb), _dot(vals, vb), 1e-12)
        outs.append(vb)
    assert torch.allclose(outs[0], outs[1], rtol=1e-12, atol=1e-12)


def _ico(level):
    from cmm_s2.core import mesh_functions as mf
    ico = mf.icosahedral_mesh(level)
    return mf.spherical_triangulation(ico.points, simplices=ico.simplices)


def _packed(mesh, jets):
    from cmm_s2 import _native as N
    from cmm_s2.core._backend import stream
    pk = torch.empty(mesh.ntri, 108, dtype=torch.float64, device="cuda")
    N.call("cmm_tri_assemble", mesh._h, jets.data_ptr(), 3, None, pk.data_ptr(), stream())
    return pk


def _jets_bar(mesh, pb):
    from cmm_s2 import _native as N
    from cmm_s2.core._backend import stream
    jb = torch.zeros(3, 4, mesh.nv, dtype=torch.float64, device="cuda")
    N.call("cmm_tri_assemble_adjoint", mesh._h, pb.data_ptr(), jb.data_ptr(), None, 0, stream())
    return jb


@pytest.mark.parametrize("scatter", ["atomics", "grouped"])
def test_compose_adjoint_fd(scatter):
    """d/dh <obar, compose(levels + h delta)> over the jets of every level."""
    from cmm_s2 import _native as N
    from cmm_s2.core._backend import stream
    from cmm_s2.core.spherical_spline import compose_packed
    from helpers import random_unit
    mesh = _ico(4)
    V = mesh.vertices
    depth = 4
    jets = [torch.from_numpy(perturbed_identity_jets(V, 40 + k, 0.05)).cuda() for k in range(depth)]
    dirs = [torch.from_numpy(perturbed_identity_jets(V, 60 + k, 1.0)).cuda() for k in range(depth)]
    for k in range(depth):  # direction: the nonlinear part only
        dirs[k] = dirs[k] - torch.from_numpy(perturbed_identity_jets(V, 60 + k, 0.0)).cuda()
    q = torch.from_numpy(random_unit(20000, 5)).cuda()
    ob = _rand(20000, 3, seed=9)

    def J(h):
        lv = [_packed(mesh, jets[k] + h * dirs[k]) for k in range(depth)]
        return _dot(compose_packed(mesh, lv, q, "N3"), ob)

    lv = [_packed(mesh, j) for j in jets]
    lb = [torch.zeros_like(p) for p in lv]
    qb = torch.empty_like(q)
    wp, wb = _work(None if scatter == "atomics" else depth * 20000, 6 * depth * mesh.ntri)
    N.call("cmm_tri_compose_adjoint", mesh._h, N.ptr_array([p.data_ptr() for p in lv]),
           N.ptr_array([p.data_ptr() for p in lb]), depth, q.data_ptr(), 20000, 3, 1, ob.data_ptr(),
           qb.data_ptr(), wp, wb, stream())
    ad = sum(_dot(_jets_bar(mesh, lb[k]), dirs[k]) for k in range(depth))
    h = 1e-6
    fd = (J(h) - J(-h)) / (2 * h)
    _close(ad, fd, 1e-6)
    # input adjoint: perturb q along tangent directions
    t = torch.cross(q, _rand(20000, 3, seed=11), dim=1)
    fdq = (_dot(compose_packed(mesh, lv, q + h * t, "N3"), ob) -
           _dot(compose_packed(mesh, lv, q - h * t, "N3"), ob)) / (2 * h)
    _close(_dot(qb, t), fdq, 1e-6)


@pytest.mark.parametrize("scatter", ["atomics", "grouped"])
def test_sample_vorticity_adjoint_fd(scatter):
    from cmm_s2 import _native as N
    from cmm_s2.core._backend import stream
    from cmm_s2.core import mesh_functions as mf, dynamical_fields as df, spherical_harmonic_tools as st
    from cmm_s2.core.spherical_simulations import VorticitySampler
    from helpers import random_unit
    L = 32
    mu = mf.structure_spherical_triangulation(L)
    w0 = st.project_onto_S12(df.lm_to_full(df.random_sh_coefficients(3, 8, L), L), L, mu)
    rot = 2 * np.pi
    smp = VorticitySampler(w0, mu, rot)
    n = 10000
    X = torch.from_numpy(random_unit(n, 4)).cuda()
    cosg = torch.from_numpy(np.repeat(np.linspace(-1, 1, 100), 100)).cuda()
    wb = _rand(n, seed=5)
    xb = torch.empty_like(X)
    pb = torch.zeros_like(w0.packed)
    wp, wsz = _work(None if scatter == "atomics" else n, 6 * mu.ntri)
    N.call("cmm_sample_vorticity_adjoint", mu._h, w0.packed.data_ptr(), X.data_ptr(), n, 3, 1, wb.data_ptr(), rot,
           xb.data_ptr(), 3, 1, pb.data_ptr(), wp, wsz, stream())
    h = 1e-6
    d = _rand(n, 3, seed=6)
    f = lambda Y: _dot(smp(Y, cos_grid=cosg, ring_len=100), wb)
    _close(_dot(xb, d), (f(X + h * d) - f(X - h * d)) / (2 * h), 1e-6)
    # coefficient adjoint: the sampled value is linear in the packed coefficients
    dp = _rand(*w0.packed.shape, seed=7)
    base = w0.packed.clone()
    w0.packed.copy_(base + dp)
    f1 = f(X)
    w0.packed.copy_(base)
    _close(_dot(pb, dp), f1 - f(X), 1e-9)


def _spline_velocity(L, seed):
    """Three velocity levels from seeded stream functions (device packed)."""
    from cmm_s2 import _native as N
    from cmm_s2.core._backend import stream
    from cmm_s2.core import mesh_functions as mf, dynamical_fields as df
    from cmm_s2.core.spherical_harmonic_tools import SHT
    mu = mf.structure_spherical_triangulation(L)
    p = SHT.get(L)
    lv, vals = [], []
    for k in range(3):
        a = df.random_sh_coefficients(seed + k, 6, L)
        w = p.full_to_packed(df.lm_to_full(a, L)) * 20
        fields = p.inverse(p.stream_operators(w, True), 12)
        v = torch.empty(12, mu.nv, dtype=torch.float64, device="cuda")
        N.call("cmm_mwss_velocity_jets", mu._h, fields.data_ptr(), v.data_ptr(), stream())
        pk = torch.empty(mu.ntri, 108, dtype=torch.float64, device="cuda")
        N.call("cmm_mwss_assemble", mu._h, v.data_ptr(), 3, None, pk.data_ptr(), stream())
        lv.append(pk)
        vals.append(v)
    return mu, lv, vals


@pytest.mark.parametrize("scatter", ["atomics", "grouped", "corners", "corners_dense"])
@pytest.mark.parametrize("integ", ["euler", "improved_euler", "rk4"])
def test_advect_stencil_adjoint_fd(integ, scatter):
    """"corners": the sparse corner-jet form, also against the packed
    level adjoints of the grouped form through the assembly transpose;
    "corners_dense": the same with ico-5 feet on an L = 16 velocity mesh
    (~40 records per triangle: groups of both reduction kernels)."""
    from cmm_s2 import _native as N
    from cmm_s2.core._backend import stream
    from cmm_s2.core.evolution_functions import _STAGES
    from cmm_s2.core.spherical_spline import lagrange_weights
    kind = {"euler": N.INT_EULER, "improved_euler": N.INT_IMPROVED_EULER, "rk4": N.INT_RK4}[integ]
    nvs = {"euler": 1, "improved_euler": 2, "rk4": 3}[integ]
    L, level = (16, 5) if scatter == "corners_dense" else (32, 3)
    mu, lv, vals = _spline_velocity(L, 70)
    mesh = _ico(level)
    nv = mesh.nv
    dt, t = 0.05, 0.1
    xs = mesh.d_spts

    def desc(levels):
        d = N.Velocity()
        d.kind = N.VEL_SPLINE
        d.nlev = nvs
        d.mesh = mu._h
        for k in range(nvs):
            d.levels[k] = levels[k].data_ptr()
        for s, ts in enumerate(_STAGES[kind](t, dt)):
            w = lagrange_weights(nvs, ts, dt, 0.0)
            for k in range(3):
                d.stage_w[s][k] = w[k]
            d.stage_t[s] = ts
        return d

    yb = _rand(3, 4 * nv, seed=3)

    def J(levels):
        ys = torch.empty(3, 4 * nv, dtype=torch.float64, device="cuda")
        N.call("cmm_advect_stencil", desc(levels), kind, dt, None, 0, xs.data_ptr(), xs.stride(0), nv, None, 0,
               ys.data_ptr(), ys.stride(0), stream())
        return _dot(ys, yb)

    def packed_form(sc):
        lb = [torch.zeros_like(p) for p in lv]
        xb = torch.empty(3, 4 * nv, dtype=torch.float64, device="cuda")
        wp, wsz = _work(None if sc == "atomics" else 16 * nv, 6 * mu.ntri)
        N.call("cmm_advect_stencil_adjoint", desc(lv), kind, dt, xs.data_ptr(), xs.stride(0), nv, yb.data_ptr(),
               yb.stride(0), N.ptr_array([p.data_ptr() for p in lb]), xb.data_ptr(), xb.stride(0), wp, wsz,
               stream())
        vbs = []
        for k in range(nvs):
            vb = torch.zeros_like(vals[k])
            N.call("cmm_mwss_assemble_adjoint", mu._h, lb[k].data_ptr(), 3, vb.data_ptr(), None, 0, stream())
            vbs.append(vb)
        return vbs, xb

    def corner_form(calls):
        # the levels start NaN-filled with clear flags; `calls` adjoint calls
        # accumulate into them
        lc = [torch.full((mu.ntri, 36), float("nan"), dtype=torch.float64, device="cuda") for _ in range(nvs)]
        lt = [torch.zeros(mu.ntri, dtype=torch.uint8, device="cuda") for _ in range(nvs)]
        xb = torch.empty(3, 4 * nv, dtype=torch.float64, device="cuda")
        wp, wsz = _work(16 * nv, mu.ntri)
        for _ in range(calls):
            N.call("cmm_advect_stencil_adjoint_corners", desc(lv), kind, dt, xs.data_ptr(), xs.stride(0), nv,
                   yb.data_ptr(), yb.stride(0), N.ptr_array([c.data_ptr() for c in lc]),
                   N.ptr_array([f.data_ptr() for f in lt]), xb.data_ptr(), xb.stride(0), wp, wsz, stream())
        vbs = []
        for k in range(nvs):
            vb = torch.zeros_like(vals[k])
            N.call("cmm_mwss_corner_jets_adjoint", mu._h, lc[k].data_ptr(), lt[k].data_ptr(), vb.data_ptr(),
                   stream())
            vbs.append(vb)
        return vbs, xb

    if scatter.startswith("corners"):
        vbs, xb = corner_form(1)
        ref, xref = packed_form("grouped")
        assert torch.equal(xb, xref)
        for a, b in zip(vbs, ref):
            assert bool(torch.isfinite(a).all())
            assert float((a - b).abs().max()) <= 1e-12 * float(b.abs().max())
        twice, _ = corner_form(2)  # accumulation: two calls give twice one
        for a, b in zip(twice, vbs):
            assert float((a - 2 * b).abs().max()) <= 1e-12 * float(b.abs().max())
        if scatter == "corners_dense":
            # 160k feet on a 1024-triangle mesh cross the C1 seams of the
            # spline within the difference step: the central difference
            # carries an O(h) error there, the form is checked against the
            # packed one above
            return
    else:
        vbs, xb = packed_form(scatter)
    dv = [_rand(12, mu.nv, seed=80 + k) for k in range(3)]
    ad = sum(_dot(vbs[k], dv[k]) for k in range(nvs))

    def perturbed(h):
        out = []
        for k in range(3):
            pk = torch.empty_like(lv[k])
            v = vals[k] + h * dv[k]
            N.call("cmm_mwss_assemble", mu._h, v.data_ptr(), 3, None, pk.data_ptr(), stream())
            out.append(pk)
        return out

    h = 1e-6
    fd = (J(perturbed(h)) - J(perturbed(-h))) / (2 * h)
    _close(ad, fd, 1e-6)
    # start-point adjoint along a smooth direction
    dx = torch.stack([torch.cos(xs[2]), torch.sin(2 * xs[0]), torch.cos(xs[1])]).contiguous()

    def Jx(x):
        ys = torch.empty(3, 4 * nv, dtype=torch.float64, device="cuda")
        N.call("cmm_advect_stencil", desc(lv), kind, dt, None, 0, x.data_ptr(), x.stride(0), nv, None, 0,
               ys.data_ptr(), ys.stride(0), stream())
        return _dot(ys, yb)

    h = 1e-6
    _close(_dot(xb, dx), (Jx((xs + h * dx).contiguous()) - Jx((xs - h * dx).contiguous())) / (2 * h), 1e-6)


def test_corner_gather_reads_flagged_triangles_only():
    """Sparse corner jets: triangles with a clear flag are never read (NaN
    there leaves the jets adjoint exactly zero); one flagged triangle adds
    the 12 rows of each of its corners into that corner's vertex."""
    from cmm_s2 import _native as N
    from cmm_s2.core._backend import stream
    from cmm_s2.core import mesh_functions as mf
    mu = mf.structure_spherical_triangulation(16)
    Tu, Nu = mu.ntri, mu.nv
    corners = torch.full((Tu, 36), float("nan"), dtype=torch.float64, device="cuda")
    flags = torch.zeros(Tu, dtype=torch.uint8, device="cuda")
    vb = torch.zeros(12, Nu, dtype=torch.float64, device="cuda")
    N.call("cmm_mwss_corner_jets_adjoint", mu._h, corners.data_ptr(), flags.data_ptr(), vb.data_ptr(), stream())
    torch.cuda.synchronize()
    assert float(vb.abs().max()) == 0.0
    t = Tu // 2 + 3  # a band triangle: three distinct non-pole vertices
    corners[t] = _rand(36, seed=9)
    flags[t] = 1
    N.call("cmm_mwss_corner_jets_adjoint", mu._h, corners.data_ptr(), flags.data_ptr(), vb.data_ptr(), stream())
    torch.cuda.synchronize()
    got = vb.cpu().numpy()
    c = corners[t].cpu().numpy().reshape(3, 12)
    touched = np.nonzero(np.abs(got).sum(axis=0))[0]
    assert len(touched) == 3
    assert sorted(tuple(got[:, v]) for v in touched) == sorted(tuple(c[i]) for i in range(3))


@pytest.mark.parametrize("scatter", ["atomics", "grouped"])
@pytest.mark.parametrize("identity", [True, False])
def test_stencil_jets_adjoint_fd(identity, scatter):
    from cmm_s2 import _native as N
    from cmm_s2.core._backend import stream
    mesh = _ico(4)
    nv = mesh.nv
    V = mesh.vertices
    eps = 1e-5
    cur_j = torch.from_numpy(perturbed_identity_jets(V, 90, 0.05)).cuda()
    dcur = torch.from_numpy(perturbed_identity_jets(V, 91, 1.0) - perturbed_identity_jets(V, 91, 0.0)).cuda()
    # feet: the stencil points moved by a smooth displacement
    sp = mesh.d_spts
    disp = 0.01 * torch.stack([torch.sin(3 * sp[1]), torch.cos(2 * sp[2]), torch.sin(sp[0])])
    ys = (sp + disp).contiguous()
    yv = (mesh.d_verts3 + 0.01 * torch.stack([torch.sin(3 * mesh.d_verts3[1]), torch.cos(2 * mesh.d_verts3[2]),
                                               torch.sin(mesh.d_verts3[0])])).contiguous()
    jb = _rand(3, 4, nv, seed=92)

    def J(feet, cj):
        jets = torch.empty(3, 4, nv, dtype=torch.float64, device="cuda")
        cur = None if identity else _packed(mesh, cj).data_ptr()
        N.call("cmm_tri_stencil_jets", mesh._h, cur, yv.data_ptr(), yv.stride(0), feet.data_ptr(), feet.stride(0),
               mesh.d_tan.data_ptr(), eps, jets.data_ptr(), None, stream())
        return _dot(jets, jb)

    pk = _packed(mesh, cur_j)
    sb = torch.empty_like(ys)
    curb = torch.zeros_like(pk)
    wp, wsz = _work(None if scatter == "atomics" else nv, 6 * mesh.ntri)
    N.call("cmm_tri_stencil_jets_adjoint", mesh._h, None if identity else pk.data_ptr(), yv.data_ptr(), yv.stride(0),
           ys.data_ptr(), ys.stride(0), mesh.d_tan.data_ptr(), eps, jb.data_ptr(), sb.data_ptr(), sb.stride(0),
           curb.data_ptr(), wp, wsz, stream())
    # a smooth direction for the feet (a per-point random one would be
    # amplified by the 1 / (4 eps) differences); the jets are quadratic in the
    # feet up to the normalised mean, so a central difference with h = 1e-5
    # has truncation error ~h^2 while its rounding (jets carry ~1e-16 / 4 eps)
    # stays ~1e-8 relative
    dfeet = torch.stack([torch.cos(sp[2]), torch.sin(2 * sp[0]), torch.cos(sp[1])]).contiguous()
    h = 1e-5
    fd = (J(ys + h * dfeet, cur_j) - J(ys - h * dfeet, cur_j)) / (2 * h)
    _close(_dot(sb, dfeet), fd, 1e-6)
    if not identity:
        h = 1e-6
        fdc = (J(ys, cur_j + h * dcur) - J(ys, cur_j - h * dcur)) / (2 * h)
        _close(_dot(_jets_bar(mesh, curb), dcur), fdc, 1e-6)


def _small_problem(seed_target=50):
    from cmm_s2.core import dynamical_fields as df, spherical_harmonic_tools as st
    from cmm_s2.core.adjoint import MisfitProblem
    L = 32
    mesh = _ico(3)
    tgt = df.random_sh_coefficients(seed_target, 8, L)
    plan = st.SHT.get(L)
    target = plan.full_to_packed(df.lm_to_full(tgt, L))
    return MisfitProblem(L, 12, 12 / L, mesh, target, n_maps=4, lmax=8)


def test_loop_gradient_fd_small():
    """dJ/dp of the whole loop (bootstrap, 3 remaps) against central
    differences along 8 random unit directions.  Tolerance: J carries
    rounding of ~1e-12 relative (the jets are 1 / (4 eps) differences of
    stencil values, eps = 1e-5), so a central difference with step h has an
    absolute noise of ~1e-12 J / h; with h = 1e-5 the truncation error is
    below 1e-5 relative.  The gradient is repeatable to ~1e-13 (atomic
    summation order)."""
    from cmm_s2.core import dynamical_fields as df
    prob = _small_problem()
    p = torch.from_numpy(prob.params.from_coeffs(df.random_sh_coefficients(500, 8, 32))).cuda()
    J, g = prob.value_and_grad(p)
    assert np.isfinite(J) and J > 0
    assert abs(prob.value(p) - J) <= 1e-14 * J
    rng = np.random.default_rng(51)
    h = 1e-5
    for _ in range(8):
        d = torch.from_numpy(rng.standard_normal(prob.params.n)).cuda()
        d /= d.norm()
        fd = (prob.value(p + h * d) - prob.value(p - h * d)) / (2 * h)
        ad = float((g * d).sum())
        assert abs(ad - fd) <= 1e-5 * abs(ad) + 2e-12 * J / h, (ad, fd)
    J2, g2 = prob.value_and_grad(p)
    assert J2 == J
    assert float((g2 - g).norm()) <= 1e-12 * float(g.norm())


def test_kept_tapes_match_recomputation():
    """Keeping the tape of every remap segment, of the last one only, or of
    none (recomputation from the checkpoints, on a second stream or on the
    same one) gives the same J and dJ/dp (to the atomic summation order,
    ~1e-13)."""
    from cmm_s2.core import dynamical_fields as df
    prob = _small_problem()
    p = torch.from_numpy(prob.params.from_coeffs(df.random_sh_coefficients(502, 8, 32))).cuda()
    J0, g0 = prob.value_and_grad(p, keep_bytes=0)  # every segment recomputed
    Js, gs = prob.value_and_grad(p, keep_bytes=0, overlap=False)  # ... on the same stream
    assert Js == J0 and float((gs - g0).norm()) <= 1e-12 * float(g0.norm())
    last = prob._segment_bytes(len(prob.ns) - 1)
    for budget in (1 << 40, last, None):
        J, g = prob.value_and_grad(p, keep_bytes=budget)
        assert J == J0
        assert float((g - g0).norm()) <= 1e-12 * float(g0.norm()), budget
    assert not prob._kept


def test_loop_value_matches_driver():
    """J recomputed from BarotropicCMM's own run of the same loop."""
    from cmm_s2.core import dynamical_fields as df, spherical_harmonic_tools as st
    from cmm_s2.core.spherical_simulations import BarotropicCMM
    prob = _small_problem()
    a = df.random_sh_coefficients(500, 8, 32)
    p = torch.from_numpy(prob.params.from_coeffs(a)).cuda()
    J = prob.value(p)
    L = 32
    w0 = st.project_onto_S12(df.lm_to_full(a, L), L, prob.mesh_u)
    sim = BarotropicCMM(L, 12, 12 / L, prob.mesh, w0, 2 * np.pi, 4)
    sim.run()
    w = sim.vorticity_on_grid(L)
    d = prob.plan.forward(w.reshape(1, L + 1, 2 * L)) - prob.target
    J2 = float((prob.cm * d * d).sum())
    assert abs(J - J2) <= 1e-10 * J2, (J, J2)


def test_autograd_function():
    from cmm_s2.core import dynamical_fields as df
    from cmm_s2.core.adjoint import misfit
    prob = _small_problem()
    p = torch.from_numpy(prob.params.from_coeffs(df.random_sh_coefficients(501, 8, 32))).cuda().requires_grad_()
    J = misfit(p, prob)
    (3.0 * J).backward()
    J2, g2 = prob.value_and_grad(p.detach())
    assert abs(float(J.detach()) - J2) <= 1e-14 * J2
    assert torch.allclose(p.grad, 3.0 * g2, rtol=1e-10, atol=0)


def test_autograd_function_batch():
    """Two misfit() calls on one problem before a single .backward() (a batch
    of initial fields summed into one loss): each call reverses its own run,
    so the gradients equal the separate value_and_grad gradients, whichever
    forward ran last."""
    from cmm_s2.core import dynamical_fields as df
    from cmm_s2.core.adjoint import misfit
    prob = _small_problem()
    ps = [torch.from_numpy(prob.params.from_coeffs(df.random_sh_coefficients(s, 8, 32))).cuda().requires_grad_()
          for s in (501, 502)]
    loss = misfit(ps[0], prob) + 2.0 * misfit(ps[1], prob)
    loss.backward()
    refs = [prob.value_and_grad(p.detach()) for p in ps]
    assert abs(float(loss.detach()) - (refs[0][0] + 2.0 * refs[1][0])) <= 1e-13 * float(loss.detach())
    assert float((refs[0][1] - refs[1][1]).norm()) > 1e-3 * float(refs[0][1].norm())  # distinct gradients
    assert torch.allclose(ps[0].grad, refs[0][1], rtol=1e-10, atol=0)
    assert torch.allclose(ps[1].grad, 2.0 * refs[1][1], rtol=1e-10, atol=0)


def test_gradient_matches_torch_autograd():
    """The hand-written adjoint against FP64 autograd through the plain-torch
    restatement of the same loop (tests/torch_cmm.py, pinned on the CPU
    against the oracle driver): L = 16, ico-2, 8 steps, remap every 3, 24
    coefficients (l <= 4).  Both differentiate the same discrete function
    for the same location decisions; the tolerance covers rounding amplified
    by the 1 / (4 eps) stencil jets."""
    import torch_cmm as TC
    from cmm_s2.core import dynamical_fields as df, spherical_harmonic_tools as st
    from cmm_s2.core.adjoint import MisfitProblem
    L, Nt, n_maps, lmax = 16, 8, 3, 4
    mesh = _ico(2)
    tgt = df.random_sh_coefficients(50, lmax, L)
    target = st.SHT.get(L).full_to_packed(df.lm_to_full(tgt, L))
    prob = MisfitProblem(L, Nt, Nt / L, mesh, target, 2 * np.pi, n_maps, lmax=lmax)
    a = df.random_sh_coefficients(500, lmax, L)
    p = prob.params.from_coeffs(a)
    J, g = prob.value_and_grad(torch.from_numpy(p).cuda())
    ref = TC.TorchCMM(L, Nt, Nt / L, mesh.vertices, mesh.simplices, tgt, 2 * np.pi, n_maps, lmax=lmax)
    pt = torch.from_numpy(p).requires_grad_()
    Jt = ref.J(pt)
    Jt.backward()
    gt = pt.grad.numpy()
    gg = g.cpu().numpy()
    assert abs(J - float(Jt.detach())) <= 1e-9 * J, (J, float(Jt.detach()))
    assert np.linalg.norm(gg - gt) <= 1e-7 * np.linalg.norm(gt), (np.linalg.norm(gg - gt), np.linalg.norm(gt))

