"""Ensembles of barotropic CMM runs on one GPU.

Members may differ in bandlimit, map mesh and initial vorticity.  Each member
owns its buffers and issues its kernels on its own HIP stream (default
priority), so the work of small members overlaps on the chip while the
host walks the members in lockstep; members share read-only tables (map
meshes, SHT plans of equal bandlimit).  A member's results do not depend on
the others: running it alone gives the same bits.
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
