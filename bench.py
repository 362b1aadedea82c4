"""Benchmark of the CMM hot path on one MI355X.

Workload (full barotropic vorticity CMM, FP64): icosahedral map mesh level 7
(163,842 vertices), MWSS velocity bandlimit L = 512, Omega = 2 pi, initial
vorticity a seeded Gaussian spherical-harmonic field (variance l^-3,
1 <= l <= 64) held as a scalar PS spline on the L = 1024 MWSS mesh, 100 RK4
steps of dt = 1/512 with a remap every 20 steps (composition depth 5), then
the composed map evaluated on the L = 2048 MWSS grid (2049 x 4096 =
8,392,704 points).

Reported: value = composed-map Mpts/s (8.39 M points through the 5 submaps,
device resident, location + evaluation, FP64), its roofline against HBM, the
CMM loop's seconds per steady-state step, and the CPU oracle timed on a
bounded sample of the same composition on the host cores.  `--steps K
--warmup W` are the timed / untimed repetitions of the composition.
"""
import argparse
import json
import os
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (ROOT, os.path.join(ROOT, "cmm-s2_amd")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

HBM_PEAK_GBS = 8000.0
# rocprofv3 counter summary of the composition kernel: its machine-readable
# lines "traffic_bytes[ico=..,Lf=..,depth=..,cells=..,tile_rows=..] = B" give
# the DRAM bytes per launch (FETCH_SIZE x read factor + WRITE_SIZE x write
# factor, the factors read off a device copy of known size in the same runs)
TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "compose_L2048_depth5_ico7_pmc.txt")
# the oracle's CPU step time at the default workload (bench.py --cpu-steps 5 on
# the GPU box's host): a line "step_s_median = S"
CPU_STEP_PROFILE = os.path.join(ROOT, "profiles", "cpu_oracle_step_L512_ico7.txt")


def _profile_value(path, key):
    """The number after "key = " in a committed profile summary, or None."""
    try:
        with open(path) as f:
            for line in f:
                if line.startswith(key + " = "):
                    return float(line[len(key) + 3:].split()[0])
    except OSError:
        pass
    return None


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--L", type=int, default=512)
    ap.add_argument("--ico", type=int, default=7)
    ap.add_argument("--Nt", type=int, default=100)
    ap.add_argument("--n-maps", type=int, default=20)
    ap.add_argument("--Lf", type=int, default=2048)
    ap.add_argument("--Lw", type=int, default=1024)
    ap.add_argument("--lmax", type=int, default=64)
    ap.add_argument("--seed", type=int, default=2)
    ap.add_argument("--cells", type=int, default=0, help="cube-map cells per face edge of the location grid (0: auto)")
    ap.add_argument("--cpu-workers", type=int, default=0,
                    help="processes of the CPU-baseline composition (0: the host CPUs this process may use, "
                         "at most 16)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=0,
                    help="instead: time K steady steps of the CPU oracle's loop at the workload (minutes per step)")
    ap.add_argument("--no-mixed", action="store_true", help="skip the FP32-Legendre comparison run")
    ap.add_argument("--ensemble", type=int, default=0,
                    help="instead: M seeded members (L in {128..512}, ico 5-7, 50 steps, remap every 10, "
                         "FP32 synthesis), one HIP stream per member, against the same members on one stream")
    ap.add_argument("--adjoint", type=int, default=0,
                    help="instead: gradient of the final-vorticity misfit for a batch of B initial fields "
                         "(L=256, ico-6, 50 steps dt=1/256, remap every 10, l <= 64) with central-difference "
                         "checks on 32 random directions")
    ap.add_argument("--synthesis", action="store_true",
                    help="time the 12-field Legendre synthesis (inverse SHT) in every arithmetic form at L=512 "
                         "and L=1024 against FP64")
    ap.add_argument("--keep-gib", type=float, default=48.0,
                    help="device memory for kept remap-segment tapes in the --adjoint comparison run")
    ap.add_argument("--fd-dirs", type=int, default=32)
    ap.add_argument("--small", action="store_true", help="quick check: L=64 ico-4 Nt=20 Lf=256")
    ap.add_argument("--save-maps", default=None, help="write the submap jets (torch.save) after the run")
    ap.add_argument("--compose-only", default=None,
                    help="load submap jets saved by --save-maps and time only the composition")
    return ap.parse_args()


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {a.gpus}); the composition and the "
              "CMM step shard by points / vertices with no collectives", file=sys.stderr)
        sys.exit(2)
    if a.cpu_steps:
        return cpu_step_bench(a)
    if a.ensemble:
        return ensemble_bench(a)
    if a.synthesis:
        return synthesis_bench(a)
    if a.adjoint:
        return adjoint_bench(a)
    if a.small:
        a.L, a.ico, a.Nt, a.n_maps, a.Lf, a.Lw = 64, 4, 20, 5, 256, 128
    import numpy as np
    import torch
    from cmm_s2.core import mesh_functions as mf, spherical_harmonic_tools as st
    from cmm_s2.core import dynamical_fields as df
    from cmm_s2.core.spherical_simulations import BarotropicCMM

    ico = mf.icosahedral_mesh(a.ico)
    pool = None
    if not a.no_cpu:
        # the CPU baseline's processes start before this one touches the GPU
        # and build their Qhull triangulations while the GPU runs
        from oracle.parallel import ComposePool
        n_w = a.cpu_workers or min(16, len(os.sched_getaffinity(0)))
        pool = ComposePool(ico.points, n_w, qhull_options=None if a.ico <= 5 else "QJ")
    torch.cuda.set_device(0)
    t0 = time.perf_counter()
    mesh = mf.spherical_triangulation(ico.points, simplices=ico.simplices, cells_per_face=a.cells)
    log(f"mesh ico-{a.ico}: {mesh.nv} vertices, {mesh.ntri} triangles, "
        f"setup {time.perf_counter() - t0:.1f}s, {mesh.stats()}")
    t0 = time.perf_counter()
    coeffs = df.random_sh_coefficients(a.seed, a.lmax, a.Lw)
    wmesh = mf.structure_spherical_triangulation(a.Lw)
    w0 = st.project_onto_S12(df.lm_to_full(coeffs, a.Lw), a.Lw, wmesh)
    torch.cuda.synchronize()
    log(f"initial vorticity spline on L={a.Lw}: {time.perf_counter() - t0:.1f}s")

    T = a.Nt / a.L
    sim = BarotropicCMM(a.L, a.Nt, T, mesh, w0, 2 * np.pi, a.n_maps)
    torch.cuda.synchronize()
    step_t, remap_t, t_boot, s_step = [], [], 0.0, None
    if a.compose_only:
        vals = torch.load(a.compose_only, weights_only=True).cuda()
        for i in range(vals.shape[0]):
            sim.submaps.assemble_coefficients(i, vals[i])
        sim.submaps.Nc = vals.shape[0]
        sim.identity = True
        torch.cuda.synchronize()
    else:
        t0 = time.perf_counter()
        sim.bootstrap()
        torch.cuda.synchronize()
        t_boot = time.perf_counter() - t0
        step_t, remap_t = run_steps(sim, pipelined=False)
        s_step_sync = statistics.median(step_t[1:] if len(step_t) > 2 else step_t)
        log(f"bootstrap {t_boot:.3f}s, synchronised steps: steady step median {s_step_sync * 1e3:.3f} ms, "
            f"remap steps {[round(x * 1e3, 3) for x in remap_t]} ms, total loop {sum(step_t) + sum(remap_t):.3f}s")
        # the same run again as run() issues it: no synchronisation between
        # steps, HIP events around every step (reset() gives the bits of a
        # first run)
        sim.reset()
        sim.bootstrap(tune=False)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        step_t, remap_t = run_steps(sim, pipelined=True)
        loop_pipe = time.perf_counter() - t0
        s_step = statistics.median(step_t[1:] if len(step_t) > 2 else step_t)
        log(f"pipelined steps: steady step median {s_step * 1e3:.3f} ms, remap steps "
            f"{[round(x * 1e3, 3) for x in remap_t]} ms, whole loop {loop_pipe:.3f}s "
            f"({loop_pipe / a.Nt * 1e3:.3f} ms per step)")
        if sim.U.sparse:
            log("velocity levels: blocks assembled on "
                f"{[round(sim.U.valid_fraction(i) * 100, 1) for i in range(3)]} % of the triangles")
        if a.save_maps:
            torch.save(sim.submaps.vals[:sim.submaps.Nc].cpu(), a.save_maps)

    # composed map on the Lf grid
    depth = sim.submaps.Nc
    grid = mf.structure_spherical_triangulation(a.Lf)
    wf_out = torch.empty((a.Lf + 1) * 2 * a.Lf, dtype=torch.float64, device="cuda")
    q = grid.sample_points()
    n = q.shape[0]
    out = torch.empty_like(q)
    levels = sim.submaps.levels(depth)
    gshape = (a.Lf + 1, 2 * a.Lf)  # ring-major MWSS grid: tiled traversal
    from cmm_s2.core.spherical_spline import compose_packed
    from cmm_s2.core.spherical_spline import tune_grid_tiles
    tiles = tune_grid_tiles(mesh, levels, q, gshape, out)
    tile_rows = min(tiles, key=tiles.get)
    log("tile rows -> ms per composition: " + ", ".join(f"{r}: {t * 1e3:.3f}" for r, t in tiles.items()))
    for _ in range(a.warmup):
        compose_packed(mesh, levels, q, "N3", out, grid=gshape)
    torch.cuda.synchronize()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(a.steps)]
    t0 = time.perf_counter()
    for e0, e1 in ev:
        e0.record()
        compose_packed(mesh, levels, q, "N3", out, grid=gshape)
        e1.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    k_ms = statistics.mean(e0.elapsed_time(e1) for e0, e1 in ev)
    ms_per = wall / a.steps * 1e3
    mpts = n / (ms_per * 1e-3) / 1e6
    alg_bytes = (48 + 216 * depth) * n
    comp = compulsory_bytes(mesh, levels, q, gshape)
    achieved = comp["bytes"] / (k_ms * 1e-3) / 1e9
    # vorticity on the fine grid (reported separately)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    wf = sim.vorticity_on_grid(a.Lf, out=wf_out, grid=grid)
    torch.cuda.synchronize()
    t_vort = time.perf_counter() - t0
    # plain copy of a known size: the HBM-counter calibration kernel
    src = torch.empty(32 * 1024 * 1024, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    log(f"compose depth {depth} on {n} points: kernel {k_ms:.3f} ms, wall {ms_per:.3f} ms/eval, "
        f"{mpts:.1f} Mpts/s; vorticity grid {t_vort * 1e3:.2f} ms; |w| max {float(wf.abs().max()):.4f}")

    area = area_distortion(sim, q)
    cons = sim.conservation_errors(1024)
    log(f"absolute vorticity of the composed map on the L=1024 grid: enstrophy rel. change "
        f"{cons['enstrophy_rel_err']:.3e}, energy rel. change {cons['energy_rel_err']:.3e} over the run")
    mixed = None
    if not a.compose_only and not a.no_mixed:
        floor = sensitivity_run(a, mesh, w0, wf, grid)
        mixed = {mode: mixed_run(a, mesh, w0, wf, grid, s_step, sim, mode, floor["vorticity_max_rel_err"])
                 for mode in _MIXED}
        mixed["fp64_initial_field_perturbed_2^-50"] = floor
    cells = int(mesh.stats()["cells_per_face"])
    traffic = _profile_value(TRAFFIC_PROFILE, f"traffic_bytes[ico={a.ico},Lf={a.Lf},depth={depth},cells={cells},"
                                              f"tile_rows={tile_rows}]")
    cpu = None
    if not a.no_cpu:
        cpu = cpu_baseline(pool, sim, q, depth, ms_per * 1e-3, s_step)

    rec = {
        "metric": "composed-map Mpts/s + s/step at L=512 ico-7; mixed vs FP64 at 1e-10 err",
        "value": round(mpts, 3), "unit": "Mpts/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": round(ms_per, 4), "higher_is_better": True, "scaling": "strong",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"barotropic CMM L={a.L} ico-{a.ico}, {a.Nt} RK4 steps dt=1/{a.L}, remap "
                               f"every {a.n_maps}; composed map (depth {depth}) on the L={a.Lf} MWSS grid",
                   "points": n, "depth": depth, "L": a.L, "ico": a.ico, "Lf": a.Lf, "Lw": a.Lw,
                   "tile_rows": tile_rows},
        "s_per_step": None if s_step is None else round(s_step, 6),
        "s_per_step_synchronised": None if s_step is None else round(s_step_sync, 6),
        "loop_s_per_step": None if s_step is None else round(loop_pipe / a.Nt, 6),
        "step_timing": "s_per_step: median over the steady steps of the GPU time between HIP events recorded "
                       "on the library's stream before and after each step, steps issued back to back as "
                       "run() issues them; s_per_step_synchronised: host time of each step with a device "
                       "synchronisation after it (includes the launch latency of the step's first kernel "
                       "and the synchronisation); loop_s_per_step: host wall time of the whole 100-step "
                       "pipelined loop (remaps included) / steps",
        "remap_step_s": [round(x, 6) for x in remap_t],
        "bootstrap_s": round(t_boot, 4), "vorticity_grid_ms": round(t_vort * 1e3, 3),
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "kernel": "k_tri_compose", "kernel_ms": round(k_ms, 4),
                     "compulsory_bytes": comp["bytes"], "compulsory": comp,
                     "traffic_GBs": None if traffic is None else round(traffic / (k_ms * 1e-3) / 1e9, 1),
                     "traffic_source": None if traffic is None else os.path.relpath(TRAFFIC_PROFILE, ROOT),
                     "gather_model_bytes": alg_bytes,
                     "gather_model_GBs": round(alg_bytes / (k_ms * 1e-3) / 1e9, 1),
                     "note": "achieved = compulsory HBM bytes per launch (points in and out, 48 B each, plus for "
                             "every level the 16-B location cell, the 128-B geometry line of each triangle and the "
                             "144-B coefficient block of each (triangle, sub-triangle) some point lands in; a lower "
                             "bound) over the HIP-event kernel time; "
                             "traffic = DRAM bytes per launch from rocprofv3 counters (traffic_source); the gather "
                             "model (48 + 216 x depth B per point) counts every lane's loads and exceeds the HBM "
                             "peak because neighbouring points share blocks in L2 / MALL; the kernel is bound by "
                             "its load-address rate"},
        "cpu_baseline": cpu,
        "mixed": mixed,
        "area_distortion": area,
        "conservation_L1024": {"enstrophy_rel_err": cons["enstrophy_rel_err"], "energy_rel_err": cons["energy_rel_err"],
                               "note": "absolute vorticity w0(X) + 2 rot cos theta(X) on the L=1024 MWSS grid against "
                                       "t = 0 (figures/data_gen_convergence.py:58-95)"},
    }
    print(json.dumps(rec), flush=True)


def ensemble_bench(a):
    """Configuration 4.  Eager lockstep runs (one HIP stream per member, and
    all members on one stream), then the whole run recorded as one HIP graph
    (forked over the member streams) and replayed.  A replay does the work of
    an eager run including its bootstrap; member steps count all Nt steps of
    every member for the full-run figures, the Nt - 2 steps after the
    bootstrap for the eager lockstep ones."""
    import gc
    import torch
    from cmm_s2.core.ensemble import Ensemble, random_members
    torch.cuda.set_device(0)
    res = {}
    for mode in ("one_stream", "streams"):
        t0 = time.perf_counter()
        builders, params = random_members(a.ensemble)
        ens = Ensemble(builders, streams=(mode == "streams"))
        setup = time.perf_counter() - t0
        n, sec = ens.timed_run()
        res[mode] = (n, sec, setup)
        log(f"ensemble {mode}: {len(params)} members, {n} member steps in {sec:.3f} s "
            f"({n / sec:.1f} member-steps/s), set-up {setup:.1f} s")
        if mode == "streams":
            break
        del ens, builders
        gc.collect()
        torch.cuda.empty_cache()
    nf, secf = ens.timed_full_run()
    log(f"ensemble eager full run (bootstrap included): {nf} member steps in {secf:.3f} s ({nf / secf:.1f}/s)")
    # every step as one launch per kernel for all members (bootstrap not timed)
    for m in ens.members:
        m.reset()
    torch.cuda.synchronize()
    batched = {}
    for parts in (1, 2, 4):
        nb, secb, per_step = ens.timed_batched_run(parts)
        batched[parts] = (nb, secb, per_step)
        log(f"ensemble batched steps, {parts} partition(s): {nb} member steps in {secb:.3f} s "
            f"({nb / secb:.1f} member-steps/s), {per_step} launches per step for {len(ens.members)} members")
        for m in ens.members:
            m.reset()
        torch.cuda.synchronize()
    nb, secb, per_step = batched[1]
    t0 = time.perf_counter()
    ens.capture()
    cap = time.perf_counter() - t0
    ens.replay()  # first replay (uploads the graph)
    torch.cuda.synchronize()
    ng, secg = ens.timed_replay(reps=3)
    log(f"ensemble graph replay: {ng} member steps in {secg:.3f} s ({ng / secg:.1f}/s), capture {cap:.1f} s")
    n, sec, _ = res["streams"]
    n1, sec1, _ = res["one_stream"]
    rec = {"metric": "ensemble member-steps/s", "value": round(ng / secg, 2), "unit": "member-steps/s",
           "n_gpus": 1, "higher_is_better": True, "data": "synthetic", "dtype": "f64 maps, fp32 synthesis",
           "config": {"workload": f"{a.ensemble} members, L uniform in {{128,192,256,384,512}}, icosahedral "
                                  f"level uniform in {{5,6,7}} (seed 4), l^-3 initial vorticity l <= L/8, "
                                  f"50 RK4 steps dt=1/L, remap every 10", "members": a.ensemble,
                      "member_params": params},
           "value_kind": "HIP-graph replay of whole runs (bootstrap included), member streams",
           "eager_full_run_value": round(nf / secf, 2),
           "eager_streams_value": round(n / sec, 2), "eager_one_stream_value": round(n1 / sec1, 2),
           "batched_value": round(nb / secb, 2), "batched_launches_per_step": per_step,
           "batched_partitions_value": {str(k): round(v[0] / v[1], 2) for k, v in batched.items()},
           "batched_partitions_launches_per_step": {str(k): v[2] for k, v in batched.items()},
           "batched_note": "every step after the bootstrap issued as one launch per kernel for all members "
                           "(member-batched grids, blocks mapped to members through a prefix table)",
           "graph_speedup_vs_eager_full_run": round(secf / (secg / 3), 4),
           "member_steps": ng, "seconds": round(secg, 4), "capture_seconds": round(cap, 2)}
    print(json.dumps(rec), flush=True)


def synthesis_bench(a):
    """The velocity projection's inverse SHT (12 fields, l^-1.5 spectrum,
    seed 7) in every Legendre form: HIP-event time per transform (20 reps,
    FFT included) and the relative max-norm difference from the FP64 form."""
    import numpy as np
    import torch
    from cmm_s2.core import spherical_harmonic_tools as st
    torch.cuda.set_device(0)
    forms = ("fp64", "fp32", "bf16x3", "bf16x4", "bf16x5", "bf16x6", "bf16x7", "i8x4", "i8x5", "i8x6", "i8x7", "i8x8")
    res = {}
    for L in (512, 1024):
        p64 = st.SHT.get(L, "fp64")
        g = torch.Generator(device="cuda").manual_seed(7)
        ls = torch.from_numpy(p64.l_of.astype(np.float64)).cuda()
        flm = torch.randn(st.nlm(L), 12, 2, dtype=torch.float64, device="cuda", generator=g)
        flm *= (1.0 + ls)[:, None, None] ** -1.5
        ref = p64.inverse(flm)
        for form in forms:
            p = st.SHT.get(L, form)
            out = p.inverse(flm)
            torch.cuda.synchronize()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            for _ in range(20):
                p.inverse(flm, out=out)
            e1.record()
            torch.cuda.synchronize()
            us = e0.elapsed_time(e1) / 20 * 1e3
            err = float((out - ref).abs().max() / ref.abs().max())
            res[f"L{L}_{form}"] = {"us": round(us, 2), "rel_err_vs_fp64": err}
            log(f"L={L} {form:7s} 12-field synthesis {us:8.1f} us, rel. max-norm diff vs fp64 {err:.2e}")
    rec = {"metric": "12-field inverse SHT time per Legendre form", "value": res["L512_bf16x6"]["us"], "unit": "us",
           "n_gpus": 1, "higher_is_better": False, "data": "synthetic", "dtype": "f64 / f32 / split-bf16",
           "config": {"workload": "inverse MWSS SHT of 12 fields (the velocity projection), L=512 and L=1024"},
           "forms": res}
    print(json.dumps(rec), flush=True)


def adjoint_bench(a):
    """Gradient of J = sum_lm |(F omega_T - omega_target)_lm|^2 (all orders)
    with respect to the initial coefficients (1 <= l <= 64) for a batch of
    seeded fields: forward and reverse times per member, and central
    differences with h = 1e-6 (and 1e-5) along seeded random unit directions
    on the first member."""
    import numpy as np
    import torch
    from cmm_s2.core import mesh_functions as mf, spherical_harmonic_tools as st, dynamical_fields as df
    from cmm_s2.core.adjoint import MisfitProblem
    torch.cuda.set_device(0)
    L, ico, Nt, n_maps, lmax = 256, 6, 50, 10, 64
    if a.small:
        L, ico, Nt, n_maps, lmax = 64, 4, 20, 5, 16
    im = mf.icosahedral_mesh(ico)
    mesh = mf.spherical_triangulation(im.points, simplices=im.simplices)
    plan = st.SHT.get(L)
    target = plan.full_to_packed(df.lm_to_full(df.random_sh_coefficients(50, lmax, L), L))
    prob = MisfitProblem(L, Nt, Nt / L, mesh, target, 2 * np.pi, n_maps, lmax=lmax)
    members = []
    ps = [torch.from_numpy(prob.params.from_coeffs(df.random_sh_coefficients(500 + k, lmax, L))).cuda()
          for k in range(a.adjoint)]
    prob.value_and_grad(ps[0])  # warm-up: plans, tile tuning, allocator
    for k, p in enumerate(ps):
        J, g = prob.value_and_grad(p)
        t = dict(prob.timings)
        members.append({"seed": 500 + k, "J": J, "grad_norm": float(g.norm()), "forward_s": round(t["forward_s"], 5),
                        "reverse_s": round(t["reverse_s"], 5),
                        "reverse_over_forward": round(t["reverse_s"] / t["forward_s"], 3)})
        log(f"member {k}: J {J:.6e} |g| {float(g.norm()):.4e} forward {t['forward_s'] * 1e3:.1f} ms "
            f"reverse {t['reverse_s'] * 1e3:.1f} ms ({t['reverse_s'] / t['forward_s']:.2f}x)")
        if k == 0:
            g0, J0 = g, J
    # the same gradients with every segment recomputed (no tape kept), and
    # with the tapes of all segments that fit a.keep_gib kept in device memory
    def variant(keep, overlap=True):
        rows = []
        for k, p in enumerate(ps):
            J, g = prob.value_and_grad(p, keep_bytes=keep, overlap=overlap)
            t = dict(prob.timings)
            rows.append({"seed": 500 + k, "forward_s": round(t["forward_s"], 5), "reverse_s": round(t["reverse_s"], 5),
                         "reverse_over_forward": round(t["reverse_s"] / t["forward_s"], 3)})
            if k == 0:
                diff = float((g - g0).norm() / g0.norm())
        fw = float(np.median([m["forward_s"] for m in rows]))
        rv = float(np.median([m["reverse_s"] for m in rows]))
        return {"forward_s_median": round(fw, 5), "reverse_s_median": round(rv, 5),
                "reverse_over_forward": round(rv / fw, 4), "grad_rel_diff_vs_default": diff, "members": rows}
    budget = int(a.keep_gib * (1 << 30))
    variants = {"recompute_every_segment": variant(0),
                "recompute_every_segment_one_stream": variant(0, overlap=False),
                f"keep_all_tapes_within_{a.keep_gib:g}_GiB": variant(budget)}
    for name, v in variants.items():
        log(f"{name}: forward {v['forward_s_median'] * 1e3:.1f} ms reverse {v['reverse_s_median'] * 1e3:.1f} ms "
            f"({v['reverse_over_forward']:.2f}x), gradient rel. diff vs default {v['grad_rel_diff_vs_default']:.1e}")
    rng = np.random.default_rng(51)
    fd = []
    for i in range(a.fd_dirs):
        d = torch.from_numpy(rng.standard_normal(prob.params.n)).cuda()
        d /= d.norm()
        ad = float((g0 * d).sum())
        row = {"adjoint": ad}
        for h in (1e-6, 1e-5):
            f = (prob.value(ps[0] + h * d) - prob.value(ps[0] - h * d)) / (2 * h)
            row[f"fd_h{h:.0e}"] = f
            row[f"rel_err_h{h:.0e}"] = abs(f - ad) / max(abs(ad), 1e-300)
        fd.append(row)
    e6 = np.array([r["rel_err_h1e-06"] for r in fd] or [np.nan])
    e5 = np.array([r["rel_err_h1e-05"] for r in fd] or [np.nan])
    ads = np.array([abs(r["adjoint"]) for r in fd] or [np.nan])
    log(f"central differences on {len(fd)} directions: rel err h=1e-6 median {np.median(e6):.2e} max {e6.max():.2e}; "
        f"h=1e-5 median {np.median(e5):.2e} max {e5.max():.2e}; |dJ.d| from {ads.min():.2e} to {ads.max():.2e}")
    fwd = float(np.median([m["forward_s"] for m in members]))
    rev = float(np.median([m["reverse_s"] for m in members]))
    rec = {"metric": "adjoint cost / forward cost (gradient of final-vorticity misfit)",
           "value": round(rev / fwd, 4), "unit": "x forward", "n_gpus": 1, "higher_is_better": False,
           "data": "synthetic", "dtype": "f64",
           "config": {"workload": f"barotropic CMM L={L} ico-{ico}, {Nt} RK4 steps dt=1/{L}, remap every {n_maps}, "
                                  f"gradient w.r.t. the initial coefficients 1 <= l <= {lmax} "
                                  f"({prob.params.n} reals), target seed 50, batch seeds 500..{499 + a.adjoint}, "
                                  "checkpoint at every remap; the last segment's tape kept, the earlier segments "
                                  "recomputed from their checkpoints on a second HIP stream under the reverse sweep",
                      "members": a.adjoint},
           "forward_s_median": round(fwd, 5), "reverse_s_median": round(rev, 5),
           "forward_ms_per_step": round(fwd / Nt * 1e3, 4), "members": members,
           "segment_gib": round(prob._segment_bytes(1) / (1 << 30), 3), "variants": variants,
           "fd_check": {"directions": len(fd), "seed": 51,
                        "rel_err_h1e-6": {"median": float(np.median(e6)), "max": float(e6.max())},
                        "rel_err_h1e-5": {"median": float(np.median(e5)), "max": float(e5.max())},
                        "abs_dir_deriv_range": [float(ads.min()), float(ads.max())], "J": J0,
                        "rows": fd[:8]}}
    print(json.dumps(rec), flush=True)


def area_distortion(sim, q):
    """Area factor |J t1 x J t2| of the composed map on the fine grid (J its
    differential on the tangent planes, t1, t2 an orthonormal tangent basis):
    exactly 1 for the incompressible flow, so its deviation measures the
    discretisation error of the maps at full size."""
    import torch
    X, J = sim.submaps.jacobian(q)
    a = torch.cross(q, torch.tensor([0.3, -0.5, 0.8], dtype=q.dtype, device=q.device).expand_as(q), dim=1)
    t1 = a / a.norm(dim=1, keepdim=True)
    t2 = torch.cross(q, t1, dim=1)
    u = (J * t1.t()[None]).sum(dim=1).t()  # J t1 per point, elementwise (no batched GEMM)
    v = (J * t2.t()[None]).sum(dim=1).t()
    f = torch.cross(u, v, dim=1).norm(dim=1)
    dev = (f - 1).abs()
    out = {"points": int(q.shape[0]), "max_abs_dev": float(dev.max()), "rms_dev": float((dev ** 2).mean().sqrt())}
    log(f"area factor of the composed map: max |f-1| {out['max_abs_dev']:.3e}, rms {out['rms_dev']:.3e}")
    return out


_MIXED = {"fp32": "fp32 matrix cores, fp64 recurrence, fp64-compensated 32-degree partial sums",
          "bf16x3": "split-bf16, 3 fixed-point digits per operand (6 digit products) on bf16 matrix cores, "
                    "exact fp32 sums of the four leading weights, fp64 flush every 192 degrees per parity",
          "bf16x5": "split-bf16, 5 fixed-point digits per operand (15 digit products) on bf16 matrix cores, "
                    "exact fp32 sums of the four leading weights, fp64 flush every 192 degrees per parity",
          "bf16x6": "split-bf16, 6 fixed-point digits per operand (21 digit products) on bf16 matrix cores, "
                    "exact fp32 sums of the four leading weights, fp64 flush every 192 degrees per parity",
          "bf16x7": "split-bf16, 7 fixed-point digits per operand (28 digit products) on bf16 matrix cores, "
                    "exact fp32 sums of the four leading weights, fp64 flush every 192 degrees per parity",
          "i8x5": "int8 Ozaki scheme, 5 int8 digits per operand (15 digit products) on int8 matrix cores, Legendre "
                  "digits tabulated once per plan, exact int32 sums per 32-degree chunk, fp64 across chunks",
          "i8x6": "int8 Ozaki scheme, 6 int8 digits per operand (21 digit products) on int8 matrix cores, Legendre "
                  "digits tabulated once per plan, exact int32 sums per 32-degree chunk, fp64 across chunks",
          "i8x7": "int8 Ozaki scheme, 7 int8 digits per operand (28 digit products) on int8 matrix cores, Legendre "
                  "digits tabulated once per plan, exact int32 sums per 32-degree chunk, fp64 across chunks"}


def sensitivity_run(a, mesh, w0, wf64, grid):
    """The FP64 run again from the initial vorticity spline with every jet
    value scaled by (1 + 2^-50): the final vorticity difference that a
    rounding-level change of the input alone produces in this run (the floor
    of any comparison of two arithmetics that round differently)."""
    import numpy as np
    from cmm_s2.core.spherical_simulations import BarotropicCMM
    from cmm_s2.core.spherical_spline import spherical_spline
    wp = spherical_spline(w0.mesh, w0.vals * (1.0 + 2.0 ** -50))
    sim = BarotropicCMM(a.L, a.Nt, a.Nt / a.L, mesh, wp, 2 * np.pi, a.n_maps)
    sim.run()
    w = sim.vorticity_on_grid(a.Lf, grid=grid)
    err = float((w - wf64).abs().max() / wf64.abs().max())
    log(f"fp64 with the initial field perturbed by 2^-50 relative: final vorticity rel. diff {err:.3e}")
    return {"perturbation": "initial vorticity jets x (1 + 2^-50)", "vorticity_max_rel_err": err}


def run_steps(sim, pipelined):
    """Step ``sim`` to its end; returns (steady-step seconds, remap-step
    seconds).  pipelined=False: a device synchronisation after every step,
    host-timed.  pipelined=True: the steps issued back to back (as run()
    does), HIP events on the library's stream between them, read after one
    synchronisation at the end."""
    import torch
    step_t, remap_t = [], []
    if not pipelined:
        while sim.count < sim.Nt:
            t0 = time.perf_counter()
            remapped = sim.step()
            torch.cuda.synchronize()
            (remap_t if remapped else step_t).append(time.perf_counter() - t0)
        return step_t, remap_t
    ev, kinds = [torch.cuda.Event(enable_timing=True)], []
    ev[0].record()
    while sim.count < sim.Nt:
        kinds.append(sim.step())
        ev.append(torch.cuda.Event(enable_timing=True))
        ev[-1].record()
    torch.cuda.synchronize()
    for k, remapped in enumerate(kinds):
        (remap_t if remapped else step_t).append(ev[k].elapsed_time(ev[k + 1]) / 1e3)
    return step_t, remap_t


def mixed_run(a, mesh, w0, wf64, grid, s64, sim64, mode, floor):
    """The same CMM run with the Legendre synthesis in a lower-precision form
    (``mode``: fp32, split-bf16 with 3 / 5 / 6 / 7 digits or int8 digit
    tables with 5 / 6 / 7 digits): seconds per steady step, the relative
    max-norm vorticity difference on the fine grid, and the relative max-norm
    difference of every intermediate output against the FP64 run: psi_lm, u
    and grad u on the grid and the velocity coefficients of the last step,
    the map positions (jet values at the vertices) of every submap.
    ``floor``: the vorticity difference of the FP64 run from a 2^-50-perturbed
    initial field (below it two arithmetics cannot be told apart)."""
    import numpy as np
    import torch
    from cmm_s2.core.spherical_simulations import BarotropicCMM
    sim = BarotropicCMM(a.L, a.Nt, a.Nt / a.L, mesh, w0, 2 * np.pi, a.n_maps, legendre=mode)
    sim.bootstrap()
    torch.cuda.synchronize()
    st, _ = run_steps(sim, pipelined=False)
    s32_sync = statistics.median(st[1:] if len(st) > 2 else st)
    sim.reset()
    sim.bootstrap(tune=False)
    torch.cuda.synchronize()
    st, _ = run_steps(sim, pipelined=True)
    s32 = statistics.median(st[1:] if len(st) > 2 else st)
    w32 = sim.vorticity_on_grid(a.Lf, grid=grid)
    err = float((w32 - wf64).abs().max() / wf64.abs().max())
    rel = lambda x, y: float((x - y).abs().max() / y.abs().max())
    sim.U.complete()
    sim64.U.complete()
    dinv = torch.tensor([0.0] + [-1.0 / (l * (l + 1)) for l in range(1, a.L)], dtype=torch.float64,
                        device="cuda")[torch.from_numpy(sim.plan.l_of).cuda()]
    per = {"psi_lm": rel(sim.w_lm[:, 0] * dinv[:, None], sim64.w_lm[:, 0] * dinv[:, None]),
           "u_grid": rel(sim.fields[0:3], sim64.fields[0:3]),
           "grad_u_grid": rel(sim.fields[3:12], sim64.fields[3:12]),
           "U_coefficients": rel(sim.U.level_packed(2), sim64.U.level_packed(2)),
           "map_positions_at_remaps": [rel(sim.submaps.vals[i][:, 0], sim64.submaps.vals[i][:, 0])
                                       for i in range(sim64.submaps.Nc)],
           "final_vorticity": err}
    log(f"{mode} Legendre, per output rel. max-norm diff vs fp64: " +
        ", ".join(f"{k} {v:.2e}" if not isinstance(v, list) else f"{k} " + "/".join(f"{x:.1e}" for x in v)
                  for k, v in per.items()))
    log(f"{mode} Legendre: steady step {s32 * 1e3:.3f} ms (fp64 {s64 * 1e3:.3f} ms), vorticity rel err {err:.3e}")
    return {"legendre": _MIXED[mode],
            "s_per_step": round(s32, 6), "fp64_s_per_step": round(s64, 6),
            "s_per_step_synchronised": round(s32_sync, 6),
            "step_speedup_vs_fp64": round(s64 / s32, 4), "vorticity_max_rel_err": err,
            "per_output_rel_err_vs_fp64_gpu": per,
            "target_rel_err": 1e-10, "meets_target": bool(err <= 1e-10 and s32 < s64),
            "rounding_floor": floor, "faster_than_fp64_at_floor": bool(err <= max(1e-10, 1.2 * floor) and s32 < s64)}


def _cube_cells(x, R):
    """Cube-map location cell of every point (the rule of the device
    locator's cube_cell: dominant axis and sign, then an R x R grid of the
    face's tangent-plane coordinates)."""
    import torch
    a = x.abs()
    k = torch.argmax(a, dim=1)  # first maximum, as the locator's strict comparisons
    m = a.gather(1, k[:, None])[:, 0]
    c = x.gather(1, k[:, None])[:, 0]
    sg = (c < 0).long()
    u = x.gather(1, ((k + 1) % 3)[:, None])[:, 0] / m
    v = x.gather(1, ((k + 2) % 3)[:, None])[:, 0] / m
    i = torch.clamp(torch.floor((u + 1.0) * 0.5 * R), 0, R - 1).long()
    j = torch.clamp(torch.floor((v + 1.0) * 0.5 * R), 0, R - 1).long()
    return ((2 * k + sg) * R + i) * R + j


def compulsory_bytes(mesh, levels, q, grid):
    """Lower bound of the composition's HBM bytes per launch: every point is
    read and written once (48 B), and for every level each location cell
    (16 B: candidate count and the first three candidates), each triangle
    geometry line (128 B: dual basis and edge norms) and each coefficient
    block (144 B: 3 components x 6 Bernstein-Bezier coefficients of one
    split sub-triangle) that some point lands in is read at least once
    (further candidates of a cell and the DRAM line granularity are not
    counted).  The levels' input points come from the composition's
    intermediates; their cell from the locator's cube-map rule, their
    triangle and sub-triangle from the device locator (first-occurrence
    argmin / argmax of the barycentrics, the split rule)."""
    import torch
    from cmm_s2.core import _backend as B
    from cmm_s2.core.spherical_spline import compose_packed
    n, depth = q.shape[0], len(levels)
    R = int(mesh.stats()["cells_per_face"])
    inter = B.empty(depth, n, 3)
    compose_packed(mesh, levels, q, "N3", B.empty(n, 3), inter=inter, grid=grid)
    blocks, tris, cells = [], [], []
    for i in range(depth):
        x = q if i == 0 else inter[i - 1]
        tri, bcc = mesh.locate(x, "N3")
        mn, mx = torch.argmin(bcc, dim=1), torch.argmax(bcc, dim=1)
        mx = torch.where(mn == mx, (mn + 1) % 3, mx)
        sub = 2 * mn + mx - (mx > mn).long()
        blocks.append(int(torch.unique(tri.long() * 6 + sub).numel()))
        tris.append(int(torch.unique(tri).numel()))
        cells.append(int(torch.unique(_cube_cells(x, R)).numel()))
    del inter
    total = 48 * n + sum(16 * c + 128 * t + 144 * b for c, t, b in zip(cells, tris, blocks))
    return {"bytes": total, "points": n, "cells_per_level": cells, "triangles_per_level": tris,
            "blocks_per_level": blocks}


def _host_info():
    cpu_model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    cpu_model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"host_cpus": os.cpu_count(), "affinity": len(os.sched_getaffinity(0)), "cpu_model": cpu_model}


def cpu_baseline(pool, sim, q, depth, gpu_compose_s, gpu_step_s):
    """The oracle's composition (scipy find_simplex + NumPy PS evaluation) of
    the same submaps on the whole fine grid, split over a spawned process
    pool as compose_maps_parallel (evolution_functions.py:158-165) splits it
    over its pool; the oracle's step time from the committed --cpu-steps
    record."""
    import numpy as np
    n_w = pool.n
    t0 = time.perf_counter()
    pool.set_levels(sim.submaps.vals[:depth].cpu().numpy())
    setup = time.perf_counter() - t0
    qs = q.cpu().numpy()
    log(f"cpu oracle: {n_w} spawned workers hold the levels after {setup:.1f} s more (assembly, not timed)")
    t0 = time.perf_counter()
    ref = pool(qs)
    dt = time.perf_counter() - t0
    pool.close()
    got = sim.submaps(q[::97].contiguous()).cpu().numpy()
    err = float(np.abs(got - ref[::97]).max())
    log(f"cpu oracle: {len(qs)} pts in {dt:.2f}s on {n_w} processes ({len(qs) / dt / 1e6:.3f} Mpts/s), "
        f"max |gpu-cpu| on every 97th point {err:.2e}")
    step = _profile_value(CPU_STEP_PROFILE, "step_s_median")
    out = {"value": round(len(qs) / dt / 1e6, 4), "unit": "Mpts/s", "cores": n_w, "kind": "port",
           "sample": f"the whole composition: {len(qs)} points of the L-grid in grid order through the {depth} "
                     f"submaps, split into {n_w} contiguous chunks over a spawned process pool "
                     "(compose_maps_parallel), scipy find_simplex + NumPy per process; set-up not timed",
           "seconds": round(dt, 3), "max_abs_diff_vs_gpu": err, **_host_info(),
           "step_s": step,
           "step_source": None if step is None else os.path.relpath(CPU_STEP_PROFILE, ROOT)}
    if step is not None and gpu_step_s:
        out["speedup_step_plus_composition"] = round((step + dt) / (gpu_step_s + gpu_compose_s), 1)
    return out


def cpu_step_bench(a):
    """The oracle's loop (oracle/driver.py, the reference's algorithm in NumPy /
    SciPy, single process; SHTs with numpy FFTs) at the default workload:
    set-up and bootstrap untimed, then the wall time of K steady steps
    (step_hook timestamps).  Minutes per step: run once and keep the record
    under profiles/."""
    import numpy as np
    from oracle import meshes as om, sht as osh, driver as odr
    from cmm_s2.core import dynamical_fields as df
    t0 = time.perf_counter()
    say = lambda msg: print(f"[{time.perf_counter() - t0:7.1f} s] {msg}", flush=True)
    pts = om.icosahedral_points(a.ico)
    oms = om.SphericalTriangulation(pts, qhull_options=None if a.ico <= 5 else "QJ")
    say(f"ico-{a.ico} triangulation (Qhull)")
    coeffs = df.random_sh_coefficients(a.seed, a.lmax, a.Lw)
    mw = om.StructuredTriangulation(a.Lw)
    say(f"L={a.Lw} MWSS mesh")
    w0 = osh.project_onto_S12(osh.lm_to_full(coeffs, a.Lw), a.Lw, mw)
    say("initial vorticity spline")
    stamps = []

    class Stop(Exception):
        pass

    def hook(count, cm, U):
        stamps.append(time.perf_counter())
        say(f"step {count}" + (f": {stamps[-1] - stamps[-2]:.2f} s" if len(stamps) > 1 else ""))
        if len(stamps) > a.cpu_steps:
            raise Stop()

    try:
        odr.euler_simulation_sphere_remapping(a.L, a.Nt, a.Nt / a.L, oms, w0, 2 * np.pi, a.n_maps,
                                              step_hook=hook, log=say)
    except Stop:
        pass
    steps = np.diff(stamps)
    rec = {"metric": "CPU oracle seconds per steady step", "value": round(float(np.median(steps)), 3), "unit": "s",
           "higher_is_better": False, "steps": [round(float(x), 3) for x in steps], "kind": "port", "cores": 1,
           "config": {"workload": f"barotropic CMM L={a.L} ico-{a.ico}, dt=1/{a.L}, remap every {a.n_maps}, "
                                  f"initial vorticity spline on L={a.Lw}", "steps_timed": len(steps),
                      "first_timed_step": 4},
           **_host_info()}
    print(json.dumps(rec), flush=True)


if __name__ == "__main__":
    main()
