"""compose_maps_parallel (cmm_s2/core/evolution_functions.py:158-165) restated
for the CPU baseline: the query points are split into n contiguous chunks
(the last one takes the remainder), and a process pool composes each chunk
with the oracle's composition (scipy find_simplex + NumPy PS evaluation,
spherical_spline.py:365-398).  Test infrastructure: only bench.py's CPU
baseline uses it.

The workers are spawned (fresh interpreters that import NumPy / SciPy only)
and build their own triangulation (Qhull, with "QJ" for large meshes) as
soon as they start, so a caller can start the pool before its own process
touches the GPU and hand over the submap jets later; each worker then
assembles its own level coefficients, so only the points and results cross
process boundaries while the pool composes, as in the reference's starmap.
"""
import numpy as np

from . import meshes
from . import splines as sp

_state = {}


def _init(vertices, qhull_options, barrier):
    _state["mesh"] = meshes.SphericalTriangulation(vertices, qhull_options=qhull_options)
    _state["barrier"] = barrier


def _set_levels(jets):
    oms = _state["mesh"]
    _state["levels"] = [sp.assemble(jets[i], oms.simplices, oms.ps_split) for i in range(len(jets))]
    # every worker takes exactly one of the n tasks: none returns before all hold one
    _state["barrier"].wait()
    return True


def _compose(q):
    return sp.compose_eval(_state["mesh"].query, _state["levels"], q)


def split(q, n):
    """evolution_functions.py:161-162 (chunks of len(q) // n, the rest last)."""
    nn = len(q) // n
    return [q[i * nn:(i + 1) * nn] for i in range(n - 1)] + [q[(n - 1) * nn:]]


class ComposePool(object):
    """A spawned pool of ``workers`` processes holding the triangulation of
    ``vertices``; ``set_levels(jets)`` gives them the composite map of the
    submap jets (depth, 3, 4, Nv); calling the pool composes points."""

    def __init__(self, vertices, workers, qhull_options="QJ"):
        import multiprocessing as mp
        import os
        ctx = mp.get_context("spawn")
        self.n = int(workers)
        # one thread per worker (the processes are the parallelism; the
        # children read the thread counts from the environment they start with)
        keys = ("OMP_NUM_THREADS", "OPENBLAS_NUM_THREADS", "MKL_NUM_THREADS")
        saved = {k: os.environ.get(k) for k in keys}
        os.environ.update({k: "1" for k in keys})
        try:
            self.pool = ctx.Pool(self.n, initializer=_init,
                                 initargs=(np.ascontiguousarray(vertices), qhull_options, ctx.Barrier(self.n)))
        finally:
            for k, v in saved.items():
                if v is None:
                    os.environ.pop(k, None)
                else:
                    os.environ[k] = v

    def set_levels(self, jets):
        jets = np.ascontiguousarray(jets)
        assert all(self.pool.map(_set_levels, [jets] * self.n, chunksize=1))

    def __call__(self, q):
        return np.vstack(self.pool.map(_compose, split(q, self.n), chunksize=1))

    def close(self):
        self.pool.close()
        self.pool.join()
