"""Checkpoint container of the barotropic CMM loop (safetensors).

The reference pickles whole Python objects (spherical_simulations.py:434-438,
453-454: ``{"vals": curr_map.vals, "Nc": submaps.Nc}`` every few steps and the
composite map at the end).  This container stores only arrays and scalars:

  tensors  submap_vals (n_maps, 3, 4, Nv)   jets of every submap slot
           curr_vals   (3, 4, Nv)           jets of the map being advected
           U_vals      (3, 3, 4, Nu)        velocity jets, time levels 0, 1, 2
           w0_vals     (4, Nw)              initial vorticity spline (if any)
           mesh_vertices (Nv, 3), mesh_simplices (Ntri, 3)  for a check
  metadata format, L, Nt, T, rot, ns, count, Nc, identity, ts, nVs, stepping,
           w0_kind (spline / analytic kind), w0_L, w0_params

Every coefficient array (submaps, current map, velocity levels) is a
deterministic function of the stored jets and is re-assembled on load.
"""
import json

import numpy as np
import torch
from safetensors.torch import load_file, save_file

from .. import _native as N
from . import _backend as B

FORMAT = "cmm-s2-checkpoint/1"


def save_state(sim, path):
    U = sim.U
    tensors = {
        "submap_vals": sim.submaps.vals.detach().cpu().contiguous(),
        "curr_vals": sim.curr.vals.detach().cpu().contiguous(),
        "U_vals": U.vals.detach().cpu().contiguous(),
        "mesh_vertices": torch.from_numpy(np.ascontiguousarray(sim.mesh.vertices, dtype=np.float64)),
        "mesh_simplices": torch.from_numpy(np.ascontiguousarray(sim.mesh.simplices, dtype=np.int64)),
    }
    meta = {
        "format": FORMAT,
        "L": sim.L, "Nt": sim.Nt, "T": sim.T, "rot": sim.rot, "ns": list(sim.ns),
        "count": sim.count, "Nc": sim.submaps.Nc, "identity": bool(sim.identity),
        "ts": [float(t) for t in U.ts], "nVs": int(U.nVs), "stepping": bool(getattr(U, "stepping", False)),
    }
    smp = sim.sampler
    if smp.kind == N.VORT_SPLINE:
        w0 = smp._keep
        tensors["w0_vals"] = w0.vals.detach().cpu().contiguous()
        meta["w0_kind"] = N.VORT_SPLINE
        meta["w0_L"] = int(w0.mesh.L)
    else:
        meta["w0_kind"] = int(smp.kind)
        meta["w0_params"] = [float(x) for x in smp.params]
    save_file(tensors, path, metadata={"cmm": json.dumps(meta)})


def read_state(path):
    """(tensors, metadata) of a checkpoint file."""
    from safetensors import safe_open
    with safe_open(path, framework="pt") as f:
        meta = json.loads(f.metadata()["cmm"])
    if meta.get("format") != FORMAT:
        raise N.CmmError(f"{path}: not a {FORMAT} file")
    return load_file(path), meta


def _vorticity_from(tensors, meta):
    from . import dynamical_fields as df
    from .mesh_functions import structure_spherical_triangulation
    from .spherical_spline import spherical_spline
    if meta["w0_kind"] == N.VORT_SPLINE:
        wm = structure_spherical_triangulation(meta["w0_L"])
        sp = spherical_spline(wm, B.to_dev(tensors["w0_vals"]))
        sp.vort = True
        return sp
    return df.DeviceVorticity("checkpointed", meta["w0_kind"], meta["w0_params"])


def load_state(cls, path, mesh, vorticity=None):
    tensors, meta = read_state(path)
    if tuple(tensors["mesh_vertices"].shape) != (mesh.nv, 3) or \
            not np.array_equal(tensors["mesh_vertices"].numpy(), np.asarray(mesh.vertices)) or \
            not np.array_equal(tensors["mesh_simplices"].numpy(), np.asarray(mesh.simplices)):
        raise N.CmmError(f"{path}: saved with a different map mesh")
    if vorticity is None:
        vorticity = _vorticity_from(tensors, meta)
    ns = meta["ns"]
    n_maps = ns[0] if len(ns) > 1 else meta["Nt"]
    sim = cls(meta["L"], meta["Nt"], meta["T"], mesh, vorticity, meta["rot"], n_maps)
    if list(sim.ns) != list(ns):
        raise N.CmmError(f"{path}: remap schedule {ns} cannot be rebuilt")
    vals = B.to_dev(tensors["submap_vals"])
    Nc = int(meta["Nc"])
    for i in range(len(ns)):
        sim.submaps.vals[i].copy_(vals[i])
    for i in range(Nc):
        sim.submaps.assemble_coefficients(i, sim.submaps.vals[i])
    sim.submaps.Nc = Nc
    U = sim.U
    uv = B.to_dev(tensors["U_vals"])
    for k in range(3):
        U.level_vals(k).copy_(uv[k])
        U.init_coeffs(k, complete=True)
    U.ts = np.array(meta["ts"], dtype=np.float64)
    U.nVs = int(meta["nVs"])
    U.stepping = bool(meta["stepping"])
    sim.count = int(meta["count"])
    sim.identity = bool(meta["identity"])
    if sim.identity:
        sim.curr = sim.map0
    else:
        from .spherical_spline import sphere_diffeomorphism
        sim.curr = sphere_diffeomorphism(mesh, B.to_dev(tensors["curr_vals"]))
    return sim
