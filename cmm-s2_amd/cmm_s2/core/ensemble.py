"""Ensembles of barotropic CMM runs on one GPU.

Members may differ in bandlimit, map mesh and initial vorticity.  Each member
owns its buffers and issues its kernels on its own HIP stream (default
priority), so the work of small members overlaps on the chip while the
host walks the members in lockstep; members share read-only tables (map
meshes, SHT plans of equal bandlimit).  A member's results do not depend on
the others: running it alone gives the same bits.

``capture()`` records one whole run of every member (bootstrap and all
steps, forked over the member streams) as a single HIP graph; ``replay()``
re-runs it from the members' initial fields with one graph launch, without
the host's per-step work.  Every kernel argument is fixed at capture (buffer
addresses, times, depths), so a replay writes the bits of the eager run; the
initial vorticity buffers may be rewritten in place between replays.
"""
import time

import numpy as np
import torch

from .spherical_simulations import BarotropicCMM


class Ensemble(object):
    def __init__(self, builders, streams=True):
        """``builders``: callables returning a BarotropicCMM; each is called
        with its member's stream current, so the member's buffers belong to
        that stream."""
        n = len(builders)
        self.own_streams = bool(streams)
        self.streams = [torch.cuda.Stream() for _ in range(n)] if streams else \
            [torch.cuda.current_stream()] * n
        self.members = []
        for b, s in zip(builders, self.streams):
            with torch.cuda.stream(s):
                self.members.append(b())
        torch.cuda.synchronize()

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
        cap = torch.cuda.Stream()
        eager = self.streams
        if not self.own_streams:  # one stream: record on the capture stream itself
            self.streams = [cap] * len(self.members)
        try:
            with torch.cuda.graph(g, stream=cap):
                for s in set(self.streams):
                    s.wait_stream(cap)
                for m, s in zip(self.members, self.streams):
                    with torch.cuda.stream(s):
                        m.bootstrap(tune=False)
                while self.step():
                    pass
                for s in set(self.streams):
                    cap.wait_stream(s)
        finally:
            self.streams = eager
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
