"""Reverse-mode gradient of a final-vorticity misfit of the barotropic CMM loop
with respect to the initial spectral coefficients.

The reference has no gradient code.  The forward function is the loop of
``euler_simulation_sphere_remapping`` (spherical_simulations.py:144-297, as
``BarotropicCMM`` runs it) with the initial vorticity a spherical-harmonic
field represented by ``project_onto_S12`` (spherical_harmonic_tools.py:
108-142) on an MWSS mesh:

    p (real coefficients, 1 <= l <= lmax)  ->  w0 spline  ->  loop  ->
    omega_T = w0(X_T) + 2 Omega (cos theta(X_T) - cos theta) on the MWSS grid
    J = sum over all (l, m), m = -l..l, of |(F omega_T)_lm - target_lm|^2

(F the forward MWSS transform; the sum over negative orders of a real field is
the weight 2 on m > 0, as the enstrophy diagnostic of
figures/data_gen_convergence.py:60-61 sums |omega_lm|^2 over all orders).

Every stage has a hand-written adjoint kernel in libcmm_s2_amd.so (the
transposes of the SHTs, spectral operators, jets, assemblies, and the
adjoints of vorticity sampling, composition, stencil projection and the RK4
feet).  The tape is kept per remap segment: the forward run stores the loop
state at every remap (velocity levels, submaps), the reverse sweep recomputes
one segment from its checkpoint with the tape on and walks it backwards, so
memory holds one segment of map and velocity coefficients at a time.

Point-location decisions are piecewise constant, so the gradient is that of
the discrete loop for the locations the forward run took; the splines are
C1, so the function is differentiable.
"""
import numpy as np
import torch

from .. import _native as N
from . import _backend as B
from .evolution_functions import _STAGES
from .mesh_functions import structure_spherical_triangulation
from .spherical_harmonic_tools import SHT, nlm, lm_index, project_onto_S12_lm
from .spherical_simulations import VorticitySampler
from .spherical_spline import compose_packed, lagrange_weights


class SpectralParams(object):
    """Real parameter vector <-> packed coefficients (nlm(L), 1, 2) of a real
    field: for l = 1..lmax, m = 0..l: Re a_l0, then Re a_lm, Im a_lm for m > 0
    (the draw order of random_sh_coefficients)."""

    def __init__(self, L, lmax):
        if not 1 <= lmax < L:
            raise N.CmmError("need 1 <= lmax < L")
        self.L, self.lmax = L, lmax
        idx = []
        for l in range(1, lmax + 1):
            for m in range(0, l + 1):
                k = lm_index(L, l, m)
                idx.append(2 * k)
                if m > 0:
                    idx.append(2 * k + 1)
        self.n = len(idx)
        self._idx_host = np.array(idx, dtype=np.int64)
        self._idx = None

    def idx(self):
        if self._idx is None:
            self._idx = torch.from_numpy(self._idx_host).to(B.device())
        return self._idx

    def to_packed(self, p, out=None):
        if out is None:
            out = B.zeros(nlm(self.L), 1, 2)
        else:
            out.zero_()
        out.view(-1)[self.idx()] = B.to_dev(p)
        return out

    def from_packed(self, flm):
        return flm.reshape(-1)[self.idx()].clone()

    def from_coeffs(self, a):
        """(L', >= lmax+1) [l, m >= 0] complex (random_sh_coefficients) -> p."""
        a = np.asarray(a)
        out = []
        for l in range(1, self.lmax + 1):
            out.append(a[l, 0].real)
            for m in range(1, l + 1):
                out += [a[l, m].real, a[l, m].imag]
        return np.array(out)


class _Velocity(object):
    """Logical velocity levels of spline_interp_velocity: ids of the VF
    outputs in level order, nVs and the time grid start ts[0]."""

    def __init__(self):
        self.ids = []
        self.nVs = 1
        self.ts0 = 0.0

    def copy(self):
        v = _Velocity()
        v.ids = list(self.ids)
        v.nVs = self.nVs
        v.ts0 = self.ts0
        return v


class MisfitProblem(object):
    """J(p) for one loop configuration; ``value`` runs the forward loop,
    ``value_and_grad`` adds the checkpointed reverse sweep."""

    EPS = 1e-5  # stencil spacing (mesh_functions.py:128)

    def __init__(self, L, Nt, T, mesh, target, rot_rate=2 * np.pi, n_maps=10, L_w=None, lmax=64):
        self.L, self.Nt, self.T, self.mesh = L, Nt, T, mesh
        self.rot = float(rot_rate)
        self.L_w = L if L_w is None else L_w
        self.params = SpectralParams(self.L_w, lmax)
        self.ns = [i for i in range(1, Nt) if i % n_maps == 0] + [Nt]
        if self.ns[0] <= 2:
            # the loop tests for a remap from step 3 on (spherical_simulations.py:
            # 263-294), so a remap at step <= 2 never fills its submap slot
            raise N.CmmError("the first remap must come after step 2 (n_maps >= 3, Nt >= 3)")
        if len(self.ns) + 1 > 8:
            raise N.CmmError("the adjoint composes at most 8 maps (remaps + current)")
        self.tspan = np.linspace(0, T, Nt, endpoint=False)
        self.dt = self.tspan[1] - self.tspan[0]
        self.mesh_u = structure_spherical_triangulation(L)
        self.mesh_w = self.mesh_u if self.L_w == L else structure_spherical_triangulation(self.L_w)
        self.plan = SHT.get(L)
        self.plan_w = SHT.get(self.L_w)
        self.samples = self.mesh_u.sample_points()
        self.grid = (L + 1, 2 * L)
        self.n = self.samples.shape[0]
        t = B.to_dev(target)
        if t.numel() != nlm(L) * 2:
            raise N.CmmError("target must be packed (nlm(L), 1, 2)")
        self.target = t.reshape(nlm(L), 1, 2)
        cm = torch.empty(nlm(L), 1, 1, dtype=torch.float64)
        for m in range(L):
            k0 = lm_index(L, m, m)
            cm[k0:k0 + (L - m)] = 1.0 if m == 0 else 2.0
        self.cm = cm.to(B.device())
        nu, tu = self.mesh_u.nv, self.mesh_u.ntri
        self._nu, self._tu = nu, tu
        self.w0 = None
        self.timings = {}
        self._sw = None  # adjoint workspace (uint8)

    def _ws(self, nbytes):
        """(pointer, bytes) of the adjoint workspace (one buffer, grown on
        demand; the calls that use it are ordered on one stream)."""
        if self._sw is None or self._sw.numel() < nbytes:
            self._sw = torch.empty(int(nbytes), dtype=torch.uint8, device=B.device())
        return self._sw.data_ptr(), self._sw.numel()

    def _scatter(self, records, key_range):
        """workspace for ``records`` grouped scatter-add records with keys
        below ``key_range``"""
        import ctypes
        b = ctypes.c_int64(0)
        N.call("cmm_scatter_workspace_bytes", int(records), int(key_range), ctypes.byref(b))
        return self._ws(b.value)

    # ---- building blocks --------------------------------------------------
    def _omega0(self, p):
        flm = self.params.to_packed(p)
        self.w0 = project_onto_S12_lm(self.plan_w, flm, self.mesh_w)
        self.sampler = VorticitySampler(self.w0, self.mesh_u, self.rot)

    def _compose(self, chain, out):
        compose_packed(self.mesh, [self.maps[i] for i in chain], self.samples, "N3", out=out, grid=self.grid)
        return out

    def _vf(self, vid, chain, rot, tape):
        """velocity from the map chain (None = the fixed sample grid)."""
        X = self.samples if chain is None else self._compose(chain, B.empty(self.n, 3))
        w = self.sampler(X, rot=rot)
        w_lm = self.plan.forward(w.reshape(1, self.L + 1, 2 * self.L))
        ops = self.plan.stream_operators(w_lm, True)
        fields = self.plan.inverse(ops, 12)
        vals = B.empty(12, self._nu)
        N.call("cmm_mwss_velocity_jets", self.mesh_u._h, fields.data_ptr(), vals.data_ptr(), B.stream())
        packed = B.empty(self._tu, 108)
        N.call("cmm_mwss_assemble", self.mesh_u._h, vals.data_ptr(), 3, None, packed.data_ptr(), B.stream())
        self.U[vid] = packed
        if tape is not None:
            tape.append(("VF", vid, chain, rot, X))

    def _desc(self, V, kind, t):
        d = N.Velocity()
        d.kind = N.VEL_SPLINE
        nl = min(V.nVs, 3)
        d.nlev = nl
        d.mesh = self.mesh_u._h
        for k in range(nl):
            d.levels[k] = self.U[V.ids[k]].data_ptr()
        for s, ts in enumerate(_STAGES[kind](t, self.dt)):
            w = lagrange_weights(V.nVs, ts, self.dt, V.ts0)
            for k in range(3):
                d.stage_w[s][k] = w[k]
            d.stage_t[s] = ts
        return d

    def _adv(self, mid, src, kind, t, V, tape):
        """map update: packed of map ``mid`` from map ``src`` (None = identity)."""
        m = self.mesh
        nv = m.nv
        desc = self._desc(V, kind, t)
        yv, ys = B.empty(3, nv), B.empty(3, 4 * nv)
        xv, xs = m.d_verts3, m.d_spts
        N.call("cmm_advect_stencil", desc, kind, float(self.dt), None if src is None else xv.data_ptr(), xv.stride(0),
               xs.data_ptr(), xs.stride(0), nv, yv.data_ptr(), yv.stride(0), ys.data_ptr(), ys.stride(0), B.stream())
        jets = B.empty(3, 4, nv)
        cur = None if src is None else self.maps[src].data_ptr()
        N.call("cmm_tri_stencil_jets", m._h, cur, yv.data_ptr(), yv.stride(0), ys.data_ptr(), ys.stride(0),
               m.d_tan.data_ptr(), self.EPS, jets.data_ptr(), None, B.stream())
        packed = B.empty(m.ntri, 108)
        N.call("cmm_tri_assemble", m._h, jets.data_ptr(), 3, None, packed.data_ptr(), B.stream())
        self.maps[mid] = packed
        if tape is not None:
            tape.append(("ADV", mid, src, kind, t, V.copy(), yv, ys))

    # ---- the loop (BarotropicCMM.bootstrap / step) ------------------------
    def _bootstrap(self, tape):
        """spherical_simulations.py:216-261."""
        V = self.V
        dt = self.dt
        self._vf(0, None, 0.0, tape)
        V.ids = [0]
        V.nVs = 1
        self._adv(0, None, N.INT_EULER, 0.0, V, tape)
        self._vf(1, [0], self.rot, tape)
        V.ids = [0, 1]
        V.nVs = 2
        self._adv(1, None, N.INT_IMPROVED_EULER, 0.0, V, tape)
        self._vf(2, [1], self.rot, tape)
        V.ids = [0, 2]
        self._adv(2, 1, N.INT_RK4, dt, V, tape)
        V.nVs = 3
        self._vf(3, [2], self.rot, tape)
        V.ids = [0, 2, 3]
        self.curr = 2
        self.count = 2
        self.next_vid, self.next_mid = 4, 3

    def _step(self, tape):
        """spherical_simulations.py:263-294."""
        V = self.V
        t = self.tspan[self.count]
        mid = self.next_mid
        self.next_mid += 1
        self._adv(mid, self.curr, N.INT_RK4, t, V, tape)
        chain = [mid] + self.submaps[::-1]
        vid = self.next_vid
        self.next_vid += 1
        self._vf(vid, chain, self.rot, tape)
        V.ids = V.ids[1:] + [vid]
        V.ts0 = float(np.float64(V.ts0) + self.dt)
        self.count += 1
        self.curr = mid
        if self.count in self.ns:
            self.submaps.append(mid)
            self.curr = None
            return True
        return False

    def _objective(self, tape):
        chain = ([] if self.curr is None else [self.curr]) + self.submaps[::-1]
        X = self._compose(chain, B.empty(self.n, 3))
        w = self.sampler(X).reshape(1, self.L + 1, 2 * self.L)
        w_lm = self.plan.forward(w)
        d = w_lm - self.target
        J = float((self.cm * d * d).sum())
        if tape is not None:
            tape.append(("OBJ", chain, X, d))
        return J

    def _state(self):
        """Loop state at a remap.  Velocity levels and maps are written once
        and never modified, so the checkpoint holds references, not copies."""
        return dict(count=self.count, curr=self.curr, V=self.V.copy(), submaps=list(self.submaps),
                    U={k: self.U[k] for k in self.V.ids}, next_vid=self.next_vid, next_mid=self.next_mid)

    def _restore(self, st):
        self.count, self.curr, self.V = st["count"], st["curr"], st["V"].copy()
        self.submaps = list(st["submaps"])
        self.U = dict(st["U"])
        self.next_vid, self.next_mid = st["next_vid"], st["next_mid"]

    def _segment_bytes(self, k):
        """Device bytes of the tape of remap segment k (its velocity levels,
        maps, sample positions and feet)."""
        steps = self.ns[k] - (self.ns[k - 1] if k > 0 else 0) + (2 if k == 0 else 0)
        per = self._tu * 864 + self.mesh.ntri * 864 + self.n * 24 + self.mesh.nv * 120
        return steps * per

    def _run(self, p, checkpoints=None, keep_bytes=0):
        """The forward loop.  With ``checkpoints`` (a list) the state at every
        remap is appended; segments whose tape fits in ``keep_bytes`` (from
        the last segment backwards) keep it, with their velocity levels and
        maps, in self._kept for a reverse sweep without recomputation."""
        self._omega0(p)
        self.V = _Velocity()
        self.U, self.maps, self.submaps = {}, {}, []
        self.curr = None
        self._kept = {}
        keep = set()
        budget = keep_bytes
        for k in range(len(self.ns) - 1, -1, -1):
            if budget >= self._segment_bytes(k):
                budget -= self._segment_bytes(k)
                keep.add(k)
        seg = 0
        tape = [] if 0 in keep else None
        self._bootstrap(tape)
        if checkpoints is not None:
            checkpoints.append(None)  # segment 0 starts from scratch
        while self.count < self.Nt:
            prev = self.curr
            remap = self._step(tape)
            if remap and tape is not None:
                self._kept[seg] = (tape, dict(self.U), dict(self.maps))
            if tape is None or remap:
                # free what no later step reads: the previous current map and
                # the velocity levels that left the time window
                if prev is not None and prev not in self.submaps:
                    self.maps.pop(prev, None)
                for k in list(self.maps):
                    if k not in self.submaps and k != self.curr:
                        del self.maps[k]
                for k in list(self.U):
                    if k not in self.V.ids:
                        del self.U[k]
            if remap:
                seg += 1
                tape = [] if seg in keep else None
                if self.count < self.Nt and checkpoints is not None:
                    checkpoints.append(self._state())

    # ---- public ------------------------------------------------------------
    def value(self, p):
        self._run(p)
        return self._objective(None)

    def forward(self, p, keep_bytes=None):
        """J(p); keeps the remap checkpoints for ``backward``.  Remap
        segments whose tape fits in ``keep_bytes`` of device memory (taken
        from the last segment backwards) keep it, so the reverse sweep skips
        their recomputation; None keeps the last segment's (the reverse
        sweep holds two segment tapes at a time anyway), 0 recomputes every
        segment from its checkpoint."""
        import time
        if keep_bytes is None:
            keep_bytes = self._segment_bytes(len(self.ns) - 1)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        self._cps = []
        self._run(p, checkpoints=self._cps, keep_bytes=keep_bytes)
        J = self._objective(None)
        self._last = self._tape_state()
        torch.cuda.synchronize()
        self.timings = {"forward_s": time.perf_counter() - t0}
        return J

    def _tape_state(self):
        """Everything ``backward`` reads of one forward run: the remap
        checkpoints, the kept segment tapes, the final maps and loop
        counters, the velocity levels and the initial-vorticity spline.  The
        tensors are referenced, not copied (they are written once), so
        holding several states lets several forward runs be reversed later,
        in any order."""
        return dict(cps=list(self._cps), kept=dict(self._kept), maps=dict(self.maps),
                    submaps=list(self.submaps), curr=self.curr, w0=self.w0, sampler=self.sampler,
                    V=self.V.copy(), U=dict(self.U), count=self.count, next_vid=self.next_vid,
                    next_mid=self.next_mid)

    def last_state(self):
        """The tape state of the last ``forward`` (pass it to ``backward``)."""
        if getattr(self, "_last", None) is None:
            raise N.CmmError("backward() needs a forward() run first")
        return self._last

    def _load_state(self, st):
        self._cps = list(st["cps"])
        self._kept = dict(st["kept"])
        self.maps, self.submaps, self.curr = dict(st["maps"]), list(st["submaps"]), st["curr"]
        self.w0, self.sampler = st["w0"], st["sampler"]
        self.V, self.U = st["V"].copy(), dict(st["U"])
        self.count, self.next_vid, self.next_mid = st["count"], st["next_vid"], st["next_mid"]

    def _record(self, k, final_maps):
        """Recompute remap segment k from its checkpoint with the tape on:
        (tape, velocity levels, maps), the form a kept segment has."""
        tape = []
        self.maps = dict(final_maps)
        cp = self._cps[k]
        if cp is None:
            self.V = _Velocity()
            self.U, self.submaps = {}, []
            self.curr = None
            self._bootstrap(tape)
        else:
            self._restore(cp)
        while self.count < self.ns[k]:
            self._step(tape)
        return tape, dict(self.U), dict(self.maps)

    @staticmethod
    def _hand_over(seg, stream):
        """Tensors of a segment recorded on another stream are used on
        ``stream``: the allocator must not reuse them before that work ends."""
        tape, U, maps = seg
        for t in list(U.values()) + list(maps.values()):
            t.record_stream(stream)
        for op in tape:
            for x in op:
                if isinstance(x, torch.Tensor):
                    x.record_stream(stream)

    def backward(self, overlap=True, state=None):
        """dJ/dp at the point of the ``forward`` run whose ``last_state()``
        is ``state`` (default: the last forward): recompute each remap
        segment from its checkpoint with the tape on and walk it backwards.
        With ``overlap`` the recomputation of segment k-1 runs on a second
        HIP stream (default priority) while segment k is reversed on the
        current one: the two are independent, and the reverse of segment k-1
        waits for its recording only.  ``state`` is not consumed."""
        import time
        self._load_state(self.last_state() if state is None else state)
        cps = self._cps
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        final_maps = {k: self.maps[k] for k in self.submaps}
        self.lam_M, self.lam_U = {}, {}
        self.lam_w = B.zeros(self.w0.packed.shape[0], self.w0.packed.shape[1])
        self._wdefer = []
        tape = []
        self._objective(tape)
        self._reverse(tape)
        kept = self._kept
        self._kept = {}
        main = torch.cuda.current_stream()
        side = None
        if overlap:
            side = getattr(self, "_side_stream", None)
            if side is None:
                side = torch.cuda.Stream()
                self._side_stream = side
        K = len(cps) - 1
        nxt = kept.pop(K) if K in kept else self._record(K, final_maps)
        for k in range(K, -1, -1):
            seg, nxt, ready = nxt, None, None
            if k > 0:
                if k - 1 in kept:
                    nxt = kept.pop(k - 1)
                elif side is not None:
                    with torch.cuda.stream(side):
                        nxt = self._record(k - 1, final_maps)
                        ready = torch.cuda.Event()
                        ready.record(side)
                    self._hand_over(nxt, main)
            tape, U, maps = seg
            self.U = U
            self.maps = dict(final_maps)
            self.maps.update(maps)
            self._reverse(tape)
            self._flush_w()
            del seg, tape, U, maps
            if ready is not None:
                main.wait_event(ready)
            elif k > 0 and nxt is None:
                nxt = self._record(k - 1, final_maps)
        self._flush_w()
        g = self._omega0_adjoint()
        torch.cuda.synchronize()
        self.timings["reverse_s"] = time.perf_counter() - t0
        self.maps, self.U, self.lam_M, self.lam_U = {}, {}, {}, {}
        self._cps, self._kept = None, {}
        return g

    def value_and_grad(self, p, keep_bytes=None, overlap=True):
        J = self.forward(p, keep_bytes)
        return J, self.backward(overlap)

    # ---- reverse sweep -----------------------------------------------------
    def _lam_map(self, mid):
        t = self.lam_M.get(mid)
        if t is None:
            t = B.zeros(self.mesh.ntri, 108)
            self.lam_M[mid] = t
        return t

    def _lam_vel(self, vid):
        """adjoint of velocity ``vid`` as sparse corner jets: (Tu, 36)
        values, valid where the (Tu,) flags are set"""
        t = self.lam_U.get(vid)
        if t is None:
            t = (B.empty(self._tu, 36), torch.zeros(self._tu, dtype=torch.uint8, device=B.device()))
            self.lam_U[vid] = t
        return t

    def _sample_and_compose_adj(self, chain, X, wbar, rot):
        n = self.n
        xbar = B.empty(n, 3) if chain else None
        # the position adjoint now; the initial-vorticity coefficient adjoint
        # of (X, wbar) later, grouped with the other steps' (_flush_w)
        if chain:
            N.call("cmm_sample_vorticity_adjoint", self.mesh_w._h, self.w0.packed.data_ptr(), X.data_ptr(), n,
                   X.stride(0), X.stride(1), wbar.data_ptr(), float(rot), xbar.data_ptr(), 3, 1, None, None, 0,
                   B.stream())
        self._wdefer.append((X, wbar))
        if chain:
            lv = N.ptr_array([self.maps[i].data_ptr() for i in chain])
            lb = N.ptr_array([self._lam_map(i).data_ptr() for i in chain])
            q = self.samples
            wp, wb = self._scatter(len(chain) * n, 6 * len(chain) * self.mesh.ntri)
            N.call("cmm_tri_compose_adjoint", self.mesh._h, lv, lb, len(chain), q.data_ptr(), n, q.stride(0),
                   q.stride(1), xbar.data_ptr(), None, wp, wb, B.stream())

    def _flush_w(self):
        """coefficient adjoint of the deferred vorticity samples: one grouped
        scatter over all their points (the per-step alternative, FP64
        atomics, cost ~50 us a step at L=256; a grouped scatter per step
        would scan the 6 Tu_w bins of the vorticity mesh every step)"""
        if not self._wdefer:
            return
        X = torch.cat([x.reshape(-1, 3) for x, _ in self._wdefer], dim=0)
        w = torch.cat([b.reshape(-1) for _, b in self._wdefer])
        self._wdefer = []
        n = X.shape[0]
        wp, wb = self._scatter(n, 6 * self.mesh_w.ntri)
        N.call("cmm_sample_vorticity_adjoint", self.mesh_w._h, self.w0.packed.data_ptr(), X.data_ptr(), n,
               X.stride(0), X.stride(1), w.data_ptr(), 0.0, None, 3, 1, self.lam_w.data_ptr(), wp, wb, B.stream())

    def _reverse(self, tape):
        plan, L = self.plan, self.L
        for op in reversed(tape):
            kind = op[0]
            if kind == "OBJ":
                _, chain, X, d = op
                lam_lm = (2.0 * self.cm * d).contiguous()
                wbar = B.empty(1, L + 1, 2 * L)
                N.call("cmm_sht_forward_adjoint", plan._h, lam_lm.data_ptr(), 1, wbar.data_ptr(),
                       plan.work(1).data_ptr(), B.stream())
                self._sample_and_compose_adj(chain, X, wbar, self.rot)
            elif kind == "VF":
                _, vid, chain, rot, X = op
                lam = self.lam_U.pop(vid, None)
                if lam is None:
                    continue
                vb = B.zeros(12, self._nu)
                N.call("cmm_mwss_corner_jets_adjoint", self.mesh_u._h, lam[0].data_ptr(), lam[1].data_ptr(),
                       vb.data_ptr(), B.stream())
                fb = B.empty(12, L + 1, 2 * L)
                N.call("cmm_mwss_velocity_jets_adjoint", self.mesh_u._h, vb.data_ptr(), fb.data_ptr(), B.stream())
                ob = B.empty(nlm(L), 12, 2)
                N.call("cmm_sht_inverse_adjoint", plan._h, fb.data_ptr(), 12, ob.data_ptr(),
                       plan.work(12).data_ptr(), B.stream())
                wlb = B.empty(nlm(L), 1, 2)
                scratch = B.empty(nlm(L), 3, 2)
                N.call("cmm_sht_stream_operators_adjoint", plan._h, ob.data_ptr(), 1, wlb.data_ptr(),
                       scratch.data_ptr(), B.stream())
                wbar = B.empty(1, L + 1, 2 * L)
                N.call("cmm_sht_forward_adjoint", plan._h, wlb.data_ptr(), 1, wbar.data_ptr(),
                       plan.work(1).data_ptr(), B.stream())
                self._sample_and_compose_adj(chain, X, wbar, rot)
            else:  # ADV
                _, mid, src, ikind, t, V, yv, ys = op
                lam = self.lam_M.pop(mid, None)
                if lam is None:
                    continue
                m = self.mesh
                nv = m.nv
                jb = B.zeros(3, 4, nv)
                wp, wb = self._ws(288 * m.ntri)
                N.call("cmm_tri_assemble_adjoint", m._h, lam.data_ptr(), jb.data_ptr(), wp, wb, B.stream())
                sb = B.empty(3, 4 * nv)
                cur = None if src is None else self.maps[src].data_ptr()
                curb = None if src is None else self._lam_map(src).data_ptr()
                # one value record per vertex (its four feet share the block),
                # grouped: no FP64 atomics
                wp, wb = self._scatter(nv, 6 * m.ntri) if cur is not None else (None, 0)
                N.call("cmm_tri_stencil_jets_adjoint", m._h, cur, yv.data_ptr(), yv.stride(0), ys.data_ptr(),
                       ys.stride(0), m.d_tan.data_ptr(), self.EPS, jb.data_ptr(), sb.data_ptr(), sb.stride(0),
                       curb, wp, wb, B.stream())
                desc = self._desc(V, ikind, t)
                lams = [self._lam_vel(V.ids[k]) for k in range(min(V.nVs, 3))]
                lcs = N.ptr_array([c.data_ptr() for c, _ in lams])
                lts = N.ptr_array([f.data_ptr() for _, f in lams])
                xs = m.d_spts
                wp, wb = self._scatter(16 * nv, self._tu)
                N.call("cmm_advect_stencil_adjoint_corners", desc, ikind, float(self.dt), xs.data_ptr(), xs.stride(0),
                       nv, sb.data_ptr(), sb.stride(0), lcs, lts, None, 0, wp, wb, B.stream())
                if mid not in self.submaps:
                    self.maps.pop(mid, None)

    def _omega0_adjoint(self):
        mw, pw = self.mesh_w, self.plan_w
        Lw = self.L_w
        vb = B.zeros(4, mw.nv)
        wp, wb = self._ws(96 * mw.ntri)
        N.call("cmm_mwss_assemble_adjoint", mw._h, self.lam_w.data_ptr(), 1, vb.data_ptr(), wp, wb, B.stream())
        fb = B.empty(4, Lw + 1, 2 * Lw)
        N.call("cmm_mwss_scalar_jets_adjoint", mw._h, vb.data_ptr(), fb.data_ptr(), B.stream())
        ob = B.empty(nlm(Lw), 4, 2)
        N.call("cmm_sht_inverse_adjoint", pw._h, fb.data_ptr(), 4, ob.data_ptr(), pw.work(4).data_ptr(), B.stream())
        flb = B.empty(nlm(Lw), 1, 2)
        N.call("cmm_sht_scalar_operators_adjoint", pw._h, ob.data_ptr(), flb.data_ptr(), B.stream())
        return self.params.from_packed(flb)


class VorticityMisfit(torch.autograd.Function):
    """J(p) as an autograd function: forward runs the loop and keeps its
    tape state (remap checkpoints, kept tapes, final maps) on ``ctx``,
    backward the recomputing reverse sweep from that state.  Several calls
    on one problem before ``.backward()`` (a batch of fields summed into one
    loss) each reverse their own run."""

    @staticmethod
    def forward(ctx, p, problem):
        ctx.problem = problem
        J = problem.forward(p.detach())
        ctx.state = problem.last_state()
        problem._last = None
        return torch.tensor(J, dtype=torch.float64, device=p.device)

    @staticmethod
    def backward(ctx, gout):
        g = ctx.problem.backward(state=ctx.state)
        ctx.state = None
        return gout * g, None


def misfit(p, problem):
    return VorticityMisfit.apply(p, problem)
