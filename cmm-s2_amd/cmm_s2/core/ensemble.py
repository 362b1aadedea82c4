"""Ensembles of barotropic CMM runs on one GPU.

Members may differ in bandlimit, map mesh and initial vorticity.  Each member
owns its buffers and issues its kernels on its own HIP stream (created with
hipStreamCreate, default priority; torch's stream pool would hand 64 members
32 shared streams), so the work of small members overlaps on the chip while
the host walks the members in lockstep; members share read-only tables (map
meshes, SHT plans of equal bandlimit).  The runtime multiplexes the streams
over the process's hardware queues (GPU_MAX_HW_QUEUES, 4 by default), which
bounds how many members run concurrently.  A member's results do not depend
on the others: running it alone gives the same bits.

``capture()`` records one whole run of every member (bootstrap and all
steps, forked over the member streams) as a single HIP graph; ``replay()``
re-runs it from the members' initial fields with one graph launch, without
the host's per-step work.  Every kernel argument is fixed at capture (buffer
addresses, times, depths), so a replay writes the bits of the eager run; the
initial vorticity buffers may be rewritten in place between replays.

``batched_step()`` advances every member by one step with ONE launch per
kernel of the step for all of them: the library records each member's
launches (cmm_batch_begin / cmm_batch_member) and issues, per kernel, one
grid whose blocks map to the members through a prefix table of their block
counts (cmm_batch_launch), on the ensemble's batch stream.  Each block runs
its member's code and arguments, so members keep the bits of their own runs.
Members whose steps differ in kind (a remap, the identity map after one, the
Legendre arithmetic, a bandlimit above 512 whose kernels run other block
sizes) are batched in separate groups.
"""
import ctypes
import time

import numpy as np
import torch

from .. import _native as N
from .spherical_simulations import BarotropicCMM

_hip = None


def _hip_lib():
    global _hip
    if _hip is None:
        _hip = ctypes.CDLL("libamdhip64.so")
        _hip.hipStreamCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
        _hip.hipStreamDestroy.argtypes = [ctypes.c_void_p]
    return _hip


class _OwnStream(object):
    """A HIP stream of its own (hipStreamCreate: default priority, not one of
    torch's pooled streams), usable wherever torch takes a stream."""

    def __init__(self):
        h = ctypes.c_void_p()
        rc = _hip_lib().hipStreamCreate(ctypes.byref(h))
        if rc != 0:
            raise N.CmmError(f"hipStreamCreate failed ({rc})")
        self.handle = h.value
        self.stream = torch.cuda.ExternalStream(self.handle)

    def destroy(self):
        if self.handle:
            _hip_lib().hipStreamDestroy(ctypes.c_void_p(self.handle))
            self.handle = None


class Ensemble(object):
    def __init__(self, builders, streams=True):
        """``builders``: callables returning a BarotropicCMM; each is called
        with its member's stream current, so the member's buffers belong to
        that stream.  ``streams`` False runs every member on one stream."""
        n = len(builders)
        self.own_streams = bool(streams)
        self._own = [_OwnStream() for _ in range(n if streams else 1)]
        # the capture stream: used by no member, waited on by all
        self._cap = _OwnStream()
        self.streams = [o.stream for o in self._own] if streams else [self._own[0].stream] * n
        self.graph = None
        self._batches = []      # one batch object and stream per partition
        self._bstreams = []
        self.batched_launches = 0
        self.partitions = None
        self.members = []
        for b, s in zip(builders, self.streams):
            with torch.cuda.stream(s):
                self.members.append(b())
        torch.cuda.synchronize()

    def close(self):
        """Release the graph and the streams (after the device is idle)."""
        torch.cuda.synchronize()
        self.graph = None
        self.members = []
        for b in self._batches:
            N.call("cmm_batch_destroy", b)
        self._batches = []
        for o in self._own + [self._cap] + self._bstreams:
            o.destroy()
        self._own = []
        self._bstreams = []

    def __del__(self):
        try:
            if self._own:
                self.close()
        except Exception:
            pass

    def bootstrap(self):
        for m, s in zip(self.members, self.streams):
            with torch.cuda.stream(s):
                m.bootstrap()

    def step(self):
        """One step of every member that has steps left; returns the number
        of member steps issued."""
        k = 0
        for m, s in zip(self.members, self.streams):
            if m.count < m.Nt:
                with torch.cuda.stream(s):
                    m.step()
                k += 1
        return k

    # bytes of device memory for one batched step's per-member arguments
    BATCH_BYTES = 8 << 20

    def set_partitions(self, n):
        """Split the members into ``n`` partitions of about equal work (cost
        of a step estimated as 2 L^2 + 5 Nv: the velocity grid and the map's
        vertex feet), each batched on its own stream, so that the hardware
        queues run the partitions' kernels concurrently."""
        if not self.own_streams:
            raise N.CmmError("batched steps need a stream per member (their work buffers are per stream)")
        n = max(1, min(int(n), len(self.members)))
        cost = [2.0 * m.L * m.L + 5.0 * m.mesh.nv for m in self.members]
        order = sorted(range(len(self.members)), key=lambda i: -cost[i])
        load = [0.0] * n
        parts = [[] for _ in range(n)]
        for i in order:  # longest first onto the least loaded partition
            p = min(range(n), key=lambda k: load[k])
            parts[p].append(i)
            load[p] += cost[i]
        self.partitions = [sorted(p) for p in parts if p]
        while len(self._batches) < len(self.partitions):
            h = N.vp()
            N.call("cmm_batch_create", int(self.BATCH_BYTES), ctypes.byref(h))
            self._batches.append(h.value)
            self._bstreams.append(_OwnStream())
        return self.partitions

    def batched_step(self):
        """One step of every member that has steps left: per partition (one
        unless ``set_partitions``), one launch per kernel for each group of
        its members whose steps are of the same kind, on the partition's batch
        stream.  Returns the number of member steps issued."""
        if self.partitions is None:
            self.set_partitions(1)
        n = ctypes.c_int64()
        launches, steps = 0, 0
        for part, b, bs in zip(self.partitions, self._batches, self._bstreams):
            groups = {}
            for i in part:
                m = self.members[i]
                if m.count < m.Nt:
                    # bandlimits up to 512 share every kernel's block size;
                    # larger ones batch only with their own bandlimit
                    shape = 0 if m.L <= 512 else m.L
                    key = ((m.count + 1) in m.ns, m.identity, m.legendre, bool(m.U.sparse), shape)
                    groups.setdefault(key, []).append(i)
            for idx in groups.values():
                N.call("cmm_batch_begin", b)
                try:
                    for k, i in enumerate(idx):
                        N.call("cmm_batch_member", b, k)
                        with torch.cuda.stream(self.streams[i]):
                            self.members[i].step()
                except BaseException:
                    N.call("cmm_batch_cancel", b)
                    raise
                N.call("cmm_batch_launch", b, bs.handle, ctypes.byref(n))
                launches += n.value
                steps += len(idx)
        self.batched_launches = launches
        return steps

    def batched_run(self, partitions=1):
        """Bootstrap on the member streams, then every step batched."""
        self.set_partitions(partitions)
        self.bootstrap()
        torch.cuda.synchronize()
        while self.batched_step():
            pass
        torch.cuda.synchronize()
        return [m.submaps for m in self.members]

    def timed_batched_run(self, partitions=1):
        """Bootstrap, then time the remaining lockstep steps issued batched:
        returns (member steps, seconds, launches of the last step)."""
        self.set_partitions(partitions)
        self.bootstrap()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        n, last = 0, 0
        while True:
            k = self.batched_step()
            if not k:
                break
            n += k
            last = self.batched_launches
        torch.cuda.synchronize()
        return n, time.perf_counter() - t0, last

    def run(self):
        self.bootstrap()
        while self.step():
            pass
        torch.cuda.synchronize()
        return [m.submaps for m in self.members]

    def capture(self):
        """Run once eagerly (allocates every buffer, tunes the tile shapes),
        then record a full run as a HIP graph.  Returns the graph."""
        self.run()
        for m in self.members:
            m.reset()
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        # one stream: record on the members' stream itself (its workspaces
        # exist from the eager run); member streams: fork from a capture
        # stream that no member uses
        cap = self._cap.stream if self.own_streams else self.streams[0]
        with torch.cuda.graph(g, stream=cap):
            for s in set(self.streams):
                if s is not cap:
                    s.wait_stream(cap)
            for m, s in zip(self.members, self.streams):
                with torch.cuda.stream(s):
                    m.bootstrap(tune=False)
            while self.step():
                pass
            for s in set(self.streams):
                if s is not cap:
                    cap.wait_stream(s)
        torch.cuda.synchronize()
        self.graph = g
        return g

    def replay(self):
        """One captured run of every member (asynchronous)."""
        self.graph.replay()

    def timed_replay(self, reps=1):
        """(member steps, seconds) of ``reps`` back-to-back replays."""
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            self.graph.replay()
        torch.cuda.synchronize()
        return reps * sum(m.Nt for m in self.members), time.perf_counter() - t0

    def timed_full_run(self):
        """(member steps, seconds) of an eager run including the bootstrap
        (the work a replay does), tile shapes already tuned."""
        for m in self.members:
            m.reset()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for m, s in zip(self.members, self.streams):
            with torch.cuda.stream(s):
                m.bootstrap(tune=False)
        while self.step():
            pass
        torch.cuda.synchronize()
        return sum(m.Nt for m in self.members), time.perf_counter() - t0

    def timed_run(self):
        """Bootstrap, then time the remaining lockstep steps: returns
        (member steps, seconds)."""
        self.bootstrap()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        n = 0
        while True:
            k = self.step()
            if not k:
                break
            n += k
        torch.cuda.synchronize()
        return n, time.perf_counter() - t0


def random_members(n, seed=4, Ls=(128, 192, 256, 384, 512), icos=(5, 6, 7), Nt=50, n_maps=10,
                   legendre="fp32", lmax_frac=8):
    """Member builders with seeded parameters: bandlimit and icosahedral level
    drawn uniformly, member k's initial vorticity a seeded (seed 100 seed + k)
    Gaussian spherical-harmonic field with variance l^-3 for 1 <= l <= L /
    lmax_frac, represented on the member's MWSS mesh; dt = 1 / L."""
    from . import dynamical_fields as df, mesh_functions as mf, spherical_harmonic_tools as st
    rng = np.random.default_rng(seed)
    Lk = rng.choice(np.array(Ls), size=n)
    ik = rng.choice(np.array(icos), size=n)
    meshes = {}
    for i in sorted(set(ik.tolist())):
        ico = mf.icosahedral_mesh(int(i))
        meshes[i] = mf.spherical_triangulation(ico.points, simplices=ico.simplices)
    builders = []
    for k in range(n):
        L, i = int(Lk[k]), int(ik[k])

        def build(L=L, i=i, k=k):
            coeffs = df.random_sh_coefficients(100 * seed + k, max(1, L // lmax_frac), L)
            w0 = st.project_onto_S12(df.lm_to_full(coeffs, L), L, mf.structure_spherical_triangulation(L))
            return BarotropicCMM(L, Nt, Nt / L, meshes[i], w0, 2 * np.pi, n_maps, legendre=legendre)
        builders.append(build)
    return builders, [(int(a), int(b)) for a, b in zip(Lk, ik)]
