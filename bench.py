#!/usr/bin/env python3
"""Throughput of the plain PBF solver on one MI355X: particle-substeps/s.

Workload: the dam break scaled to a 2^20 particle budget (Simulation's fit
gives 1,028,772 particles), one scene, 2 substeps x 4 constraint iterations
per step(1/60), Simulation's default fluid effects and the five container
planes. Inputs are resident in HBM before the timed region; every step is a
full solver step (sort, neighbour list, surface tension, 4 x (lambda, delta +
walls), advect, XSPH + vorticity, friction).

Prints one JSON line. `roofline` is the dominant kernel's algorithmic bytes
per launch over its average duration (HIP events on the library stream);
`roofline_substep` applies the same byte model to the whole substep. The
per-kernel durations come from a twin of the scene stepped launch by launch
over the same steps (the solver is deterministic, so the twin passes through
the same states as the timed scene).
`cpu_baseline` times the CPU oracle (the FP32 restatement of the reference
schedule, the host's cores) on steps of the same scene; `cpu_c1` times the
5,000-budget dam break (4,998 particles, 200 steps) on one thread and on all
cores. `sort` times the isolated 4,194,304-point hash-grid sort.
"""
import argparse
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "threejs-particle-fluids_amd"))

HBM_PEAK_GBS = 8000.0

# Algorithmic bytes per particle per launch of each kernel: compulsory state
# traffic between neighbour barriers (FP32 vec3 = 12 B, scalar = 4 B;
# neighbour gathers, neighbour lists and collider multipliers count 0;
# per-bucket arrays 4 B a touch).
KERNEL_BYTES = {
    "hash_predict": 32,        # r x, v; w key, rank
    "place": 24,               # r key, rank, id; w src, id, key
    "rank_permute": 88,        # r key, id, src + x, v, rho; w x, v, x*, rho, id, slot
    "neighbors": 24,           # r x*; w colour normal
    "impulses": 64,            # r x*, n, rho, v; w v, x*
    "lambda": 20,              # r x*; w rho, lambda
    "delta_walls": 28,         # r x*, lambda; w x*  (+24 advect on the last iteration)
    "delta": 28,
    "velocity": 52,            # r x, v; w omega, |omega|, dv
    "vorticity_friction": 64,  # r x, |omega|, omega, dv, v; w v
    "vorticity": 64,
}
SUBSTEP_BYTES_PER_PARTICLE = lambda iters: 336 + 48 * iters  # noqa: E731

# Measured HBM bytes per particle per launch (FETCH_SIZE + WRITE_SIZE, each
# calibrated against a copy of known size) of the 1M dam break over the
# bench's window (steps 5-25), from the rocprofv3 counter summary
# profiles/dam_break_1m_hbm_traffic.md.
HBM_TRAFFIC_B_PER_PARTICLE = {
    "neighbors": 667.2,
    "impulses": 402.2,
    "lambda": 192.5,
    "delta_walls": 224.5,
    "velocity": 213.8,
    "vorticity_friction": 264.9,
    "rank_permute": 238.7,
    "hash_predict": 59.6,
    "impulse_apply": 116.1,
    "place": 29.8,
}
SCAN_BYTES_PER_BUCKET = 12

# VALU instructions (wave64 SQ_INSTS_VALU) per particle per launch of the
# 1M dam break's kernels over the bench's window, from the rocprofv3 counter
# summary profiles/dam_break_1m_sq_counters.md. The VALU peak is the MI355X's
# issue rate with packed FP32 off: 256 CUs x 4 SIMDs x 2.4 GHz, one wave64
# instruction per SIMD every 4 cycles.
VALU_INSTS_PER_PARTICLE = {
    "neighbors": 179.4,
    "impulses": 73.0,
    "lambda": 43.7,
    "delta_walls": 55.1,
    "velocity": 63.2,
    "vorticity_friction": 42.6,
    "rank_permute": 3.9,
    "hash_predict": 2.8,
}
VALU_PEAK_WAVE_INSTS = 256 * 4 * 2.4e9 / 4
# The whole substep's VALU instructions per particle, from the same summary:
# the kernels once per substep (neighbour build 179.4, surface tension 73.0
# and its apply 0.5, velocity 63.2, vorticity 42.6, the sort 7.0) and the
# density and correction passes once per iteration (43.7 + 55.1).
VALU_SUBSTEP_INSTS_PER_PARTICLE = lambda iters: 365.7 + 98.8 * iters  # noqa: E731

# The isolated sort: 4,194,304 positions, cell 0.01, 2^22 buckets. Algorithmic
# bytes per sort 32 N + 12 H: per point read x (16), write and read its key
# (8), write its sorted key and its permutation entry (8); per bucket clear the
# count, read it for the scan, write the offset (12).
SORT_N = 4_194_304
SORT_H = 1 << 22
SORT_CELL = 0.01
SORT_BYTES = 32 * SORT_N + 12 * SORT_H


def host_threads():
    """Host threads for the CPU baseline: the cores this process may use,
    at most 16 (a one-GPU box's share)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "host"


def run_sort(args, ctx):
    """The hash-grid sort in isolation (HashGrid.rebuild over a
    standalone grid): positions resident in HBM, outputs the keys, counts,
    offsets, sorted keys and permutation; both seeded distributions."""
    import particle_fluids as pf
    from particle_fluids import scenes

    out = {"workload": f"hash-grid sort (standalone rebuild), {SORT_N} positions, cell {SORT_CELL}, {SORT_H} buckets",
           "particles": SORT_N, "hash_table_size": SORT_H, "bytes_per_sort": SORT_BYTES, "reps": args.sort_reps}
    for dist in ("uniform", "blobs"):
        pos = scenes.sort_points(SORT_N, dist)
        g = pf.HashGrid.standalone(SORT_N, SORT_CELL, SORT_H)
        g.setSortedPositions(False)
        g.uploadPositions(pos)
        for _ in range(3):
            g.rebuild()
        ctx.synchronize()
        ctx.timer_begin()
        for _ in range(args.sort_reps):
            g.rebuild()
        ms = ctx.timer_end() / args.sort_reps
        ctx.set_profiling(True)
        ctx.profile_reset()
        for _ in range(5):
            g.rebuild()
        stats = ctx.profile()
        ctx.set_profiling(False)
        ctx.profile_reset()
        over = g.readbackOverflow()
        g.dispose()
        gbs = SORT_BYTES / (ms / 1000.0) / 1e9
        out[dist] = {"us_per_sort": round(ms * 1000.0, 2),
                     "roofline": {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                  "frac": round(gbs / HBM_PEAK_GBS, 4)},
                     "overflow": bool(over), "kernels": kernel_table(stats)}
        print(f"[bench] sort {dist}: {out[dist]['us_per_sort']} us, {gbs:.0f} GB/s", file=sys.stderr, flush=True)
    return out


def oracle_dam_break(orc, pos, vel, r, substeps, iterations, H):
    from particle_fluids import scenes

    n = len(pos)
    o = orc.Scene(n, r, "f32")
    o.set_particles(pos, vel)
    o.set_loop(substeps, iterations, (0.0, -9.81, 0.0), 4 * r, H)
    o.add_fluid(0, n, 1000.0, 2 * r, 4 * r, 1e-4, viscosity=0.01, vorticity=0.02, surface_tension=0.1, adhesion=0.1)
    mn, mx = scenes.CONTAINER[:3], scenes.CONTAINER[3:]
    for nrm, which in scenes.PLANES:
        o.add_plane(nrm, mn if which == "min" else mx)
    return o


def run_cpu_c1(args):
    """The reference's CPU-runnable dam break: a 5,000 budget (Simulation's
    fit: 4,998 particles), +-1% jitter (seed 0xC0FFEE), 2 substeps x 4
    iterations, 200 steps of 1/60, on the FP32 oracle: one thread, then the
    host's cores. A restatement of the reference (which has no CPU solver);
    baseline only."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as orc
    from particle_fluids import scenes

    pos, r = scenes.dam_break_layout(5000)
    pos = scenes.jitter(pos, 2 * r)
    n = len(pos)
    H = 1
    while H < 2 * n:
        H <<= 1
    out = {"workload": f"dam break, {n} particles (budget 5000), {args.substeps} substeps x {args.iterations} "
                       f"iterations, dt 1/60, {args.cpu_c1_steps} steps, FP32 C oracle",
           "particles": n, "steps": args.cpu_c1_steps, "cpu": cpu_model(), "kind": "port"}
    for label, threads in (("one_thread", 1), ("all_cores", host_threads())):
        orc.set_threads(threads)
        o = oracle_dam_break(orc, pos, None, r, args.substeps, args.iterations, H)
        t = time.perf_counter()
        for _ in range(args.cpu_c1_steps):
            o.step(1.0 / 60.0)
        sec = time.perf_counter() - t
        out[label] = {"threads": threads, "seconds": round(sec, 3),
                      "particle_substeps_per_s": round(n * args.substeps * args.cpu_c1_steps / sec, 1)}
        print(f"[bench] C1 on the CPU, {threads} thread(s): {sec:.2f} s", file=sys.stderr, flush=True)
    orc.set_threads(0)
    return out


def run_batched(args, ctx):
    """256 independent scenes, particle budgets log-uniform in [5,000, 200,000];
    every other scene adds 20 soft cubes of 2,197 particles (global shape
    matching, stiffness uniform in [0.1, 1.0]) with particle contacts on, as
    Simulation sets them for scenes with soft bodies. Times the same scenes stepped
    one loop per launch (each loop's own HIP graph) and as one batch (one HIP
    graph over every scene), both with inputs resident in HBM."""
    import particle_fluids as pf
    from particle_fluids import scenes

    dt = 1.0 / 60.0
    counts = scenes.log_uniform_counts(args.batch_scenes, 5000, 200000, args.batch_seed)
    t0 = time.perf_counter()
    built = []
    for i, c in enumerate(counts):
        built.append(scenes.build_mixed_scene(int(c), 20 if i % 2 else 0, seed=args.batch_seed + i,
                                              substeps=args.substeps, iterations=args.iterations, use_graph=True))
    loops = [b["loop"] for b in built]
    n_total = sum(b["ps"].capacity for b in built)
    n_fluid = sum(b["layout"]["fluid_count"] for b in built)
    print(f"[bench] batched: {len(built)} scenes, {n_total} particles ({n_fluid} fluid), built in "
          f"{time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
    batch = pf.Batch(loops)
    batch.useGraph(True)
    for _ in range(args.batch_warmup):
        batch.step(dt)
    tuned = None
    if not args.no_autotune:
        # the tile-union capacity, then the launch shape (one workgroup per
        # tile, or the persistent form at 4, 8 or 16 workgroups per CU) at it
        best, times = batch.autotune(dt, steps=2)
        shape, shape_times = batch.autotuneLaunch(dt, candidates=(0, 1024, 2048, 4096), steps=2)
        tuned = {"tile_union_capacity": best, "ms_per_step": {str(k): round(v, 3) for k, v in times.items()},
                 "launch_shape": shape,
                 "launch_ms_per_step": {("per_tile" if k == 0 else f"persistent_{k}"): round(v, 3)
                                        for k, v in shape_times.items()}}
        print(f"[bench] batched autotune: {tuned}", file=sys.stderr, flush=True)
    for L in loops:  # the plain form's graphs, captured before timing
        L.step(dt)
    ctx.synchronize()

    def timed(fn):
        ctx.synchronize()
        t = time.perf_counter()
        fn()
        ctx.synchronize()
        return time.perf_counter() - t

    def plain_step():
        for L in loops:
            L.step(dt)

    # Alternate one plain round (every scene stepped by its own launches) and
    # one batched step, so both forms see nearly the same scene states: the
    # cost of a step changes quickly while the dam breaks collapse.
    t_plain = t_batch = 0.0
    for _ in range(args.batch_steps):
        t_plain += timed(plain_step)
        t_batch += timed(lambda: batch.step(dt))
    # per-kernel breakdown of one batched step, launched kernel by kernel
    batch.useGraph(False)
    ctx.set_profiling(True)
    ctx.profile_reset()
    batch.step(dt)
    stats = ctx.profile()
    ctx.set_profiling(False)
    total_ms = sum(t for _, t in stats.values())
    kernels = {k: {"launches": v[0], "ms": round(v[1], 3), "share": round(v[1] / total_ms, 4)}
               for k, v in sorted(stats.items(), key=lambda kv: -kv[1][1])}
    work = n_total * args.substeps * args.batch_steps
    out = {
        "workload": f"{len(built)} scenes, budgets log-uniform in [5000, 200000] (seed {args.batch_seed}), "
                    f"every other scene + 20 soft cubes of 2197 particles with particle contacts, {args.substeps} substeps x "
                    f"{args.iterations} iterations, dt 1/60",
        "scenes": len(built), "particles": n_total, "fluid_particles": n_fluid,
        "steps": args.batch_steps, "timing": "alternating plain rounds and batched steps",
        "plain": {"particle_substeps_per_s": round(work / t_plain, 1),
                  "scene_steps_per_s": round(len(built) * args.batch_steps / t_plain, 2),
                  "ms_per_round": round(t_plain / args.batch_steps * 1e3, 3)},
        "batched": {"particle_substeps_per_s": round(work / t_batch, 1),
                    "scene_steps_per_s": round(len(built) * args.batch_steps / t_batch, 2),
                    "ms_per_step": round(t_batch / args.batch_steps * 1e3, 3)},
        "speedup": round(t_plain / t_batch, 3),
        "autotune": tuned,
        "kernels_one_step": kernels,
    }
    print(f"[bench] batched: plain {out['plain']}, batched {out['batched']}, x{out['speedup']}",
          file=sys.stderr, flush=True)
    batch.dispose()
    for b in built:
        scenes.dispose_scene(b)
    return out


# Algorithmic bytes of one surface-field update: per voxel the clear (20),
# the walls (8) and the three blur passes (40, 40, 44: int or float4
# accumulators plus the solid channel in and out; the last reads the walls and
# writes the float4 field); per particle the splat's reads (x, v, id: 36).
FIELD_BYTES_PER_VOXEL = 152
FIELD_BYTES_PER_PARTICLE = 36
# Viscosity solve per fluid particle: seed and apply (48 each) and 48 per
# Jacobi sweep (read v0 and the previous iterate's own value, write the next;
# neighbour gathers count 0).
VISCOSITY_BYTES = lambda sweeps: 48 * (sweeps + 2)  # noqa: E731


def kernel_table(stats):
    total = sum(t for _, t in stats.values()) or 1.0
    return {k: {"launches": v[0], "avg_us": round(v[1] / v[0] * 1000.0, 3), "share": round(v[1] / total, 4)}
            for k, v in sorted(stats.items(), key=lambda kv: -kv[1][1])}


def run_surface_field(args, ctx, sc):
    """SurfaceField over the stepped 1M dam break: the renderer's per-frame
    update at its largest default grid (defaultVoxelBudget 1.2 M voxels for
    50,000 particles or more), wetted container walls, motionStretch 0.02."""
    import particle_fluids as pf
    from particle_fluids import scenes

    mn, mx = scenes.CONTAINER[:3], scenes.CONTAINER[3:]
    field = pf.SurfaceField(sc["fluid"], (mn, mx), motionStretch=0.02, wet=True)
    nx, ny, nz = field.resolution
    voxels = nx * ny * nz
    n = sc["ps"].capacity
    for _ in range(3):
        field.update()
    ctx.synchronize()
    ctx.timer_begin()
    for _ in range(args.field_updates):
        field.update()
    ms = ctx.timer_end() / args.field_updates
    ctx.set_profiling(True)
    ctx.profile_reset()
    for _ in range(5):
        field.update()
    stats = ctx.profile()
    ctx.set_profiling(False)
    ctx.profile_reset()
    field.dispose()
    bytes_ = FIELD_BYTES_PER_VOXEL * voxels + FIELD_BYTES_PER_PARTICLE * n
    achieved = bytes_ / (ms / 1000.0) / 1e9
    out = {"workload": f"SurfaceField update over the {n}-particle dam break: {nx}x{ny}x{nz} voxels "
                       f"(defaultVoxelBudget), motionStretch 0.02, container walls wetted",
           "voxels": voxels, "particles": n, "updates": args.field_updates, "ms_per_update": round(ms, 4),
           "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": round(achieved / HBM_PEAK_GBS, 4), "bytes_per_update": bytes_},
           "kernels": kernel_table(stats)}
    print(f"[bench] surface field: {out['ms_per_update']} ms per update, {voxels} voxels", file=sys.stderr,
          flush=True)
    return out


def run_viscosity(args, ctx, dt):
    """The 1M dam break with a ViscositySolver after the fluid (a thick
    liquid, Simulation.addFluid({thickness}) at honey's 20 m^2/s, 12 Jacobi
    sweeps per substep), as a HIP graph."""
    import particle_fluids as pf
    from particle_fluids import scenes

    pos, r = scenes.dam_break_layout(args.budget)
    pos = scenes.jitter(pos, 2 * r)
    n = len(pos)
    ps = pf.ParticleSystem(n, r)
    ps.uploadParticles(pos)
    fluid = pf.FluidSystem(ps, **scenes.FLUID_DEFAULTS)
    solver = pf.ViscositySolver(fluid, 20.0, iterations=args.viscosity_sweeps)
    walls = pf.PrimitiveSet(ps)
    mn, mx = scenes.CONTAINER[:3], scenes.CONTAINER[3:]
    for nrm, which in scenes.PLANES:
        walls.addPlane(nrm, mn if which == "min" else mx)
    loop = pf.SimLoop(ps, substeps=args.substeps, iterations=args.iterations, materials=[fluid, solver],
                      colliders=[walls])
    loop.useGraph(True)
    for _ in range(args.warmup):
        loop.step(dt)
    ctx.synchronize()
    ctx.timer_begin()
    t0 = time.perf_counter()
    for _ in range(args.viscosity_steps):
        loop.step(dt)
    gpu_ms = ctx.timer_end()
    seconds = max(time.perf_counter() - t0, gpu_ms / 1000.0)
    loop.useGraph(False)
    ctx.set_profiling(True)
    ctx.profile_reset()
    for _ in range(3):
        loop.step(dt)
    stats = ctx.profile()
    ctx.set_profiling(False)
    ctx.profile_reset()
    for obj in (loop, solver, fluid, walls, ps):
        obj.dispose()
    launches, total = stats.get("viscosity_solve", (0, 0.0))
    out = {"workload": f"dam break, {n} particles, ViscositySolver 20 m^2/s with {args.viscosity_sweeps} sweeps, "
                       f"{args.substeps} substeps x {args.iterations} iterations, dt 1/60",
           "particles": n, "steps": args.viscosity_steps,
           "particle_substeps_per_s": round(n * args.substeps * args.viscosity_steps / seconds, 1),
           "ms_per_step": round(seconds / args.viscosity_steps * 1000.0, 4), "kernels": kernel_table(stats)}
    if launches:
        avg_s = total / launches / 1000.0
        b = VISCOSITY_BYTES(args.viscosity_sweeps) * n
        out["roofline"] = {"bound": "hbm", "kernel": "viscosity_solve", "achieved": round(b / avg_s / 1e9, 1),
                           "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(b / avg_s / 1e9 / HBM_PEAK_GBS, 4),
                           "bytes_per_launch": b, "avg_launch_us": round(avg_s * 1e6, 3)}
    print(f"[bench] viscosity: {out['particle_substeps_per_s']:.4e} particle-substeps/s", file=sys.stderr, flush=True)
    return out


def process_mappings():
    """Memory mappings this process holds (each device chunk of an evictable
    scene is one or more), or -1 where /proc is unavailable."""
    try:
        with open("/proc/self/maps") as f:
            return sum(1 for _ in f)
    except OSError:
        return -1


def device_free_gib():
    """Free device memory (hipMemGetInfo), GiB."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    free, total = ctypes.c_size_t(), ctypes.c_size_t()
    if hip.hipMemGetInfo(ctypes.byref(free), ctypes.byref(total)) != 0:
        return -1.0
    return free.value / 2 ** 30


def max_mappings():
    try:
        with open("/proc/sys/vm/max_map_count") as f:
            return int(f.read().strip())
    except (OSError, ValueError):
        return -1


def run_served(args, ctx, tile_union=None):
    """1,000 live scenes (fluid budgets log-uniform in [5,000, 500,000], 0-30
    soft cubes each, particle contacts on in scenes with cubes, all from a
    seed), resident in HBM under a 200 GB budget,
    served by the request loop with continuous batching. A Poisson stream of
    step(1/60) requests at --serve-rate req/s picks scenes uniformly; reports
    throughput and p50/p99 latency (submission to completion), then the
    capacity with every request queued at once, then the plain form (each
    request as its own loop step, HIP graph per loop) on the same kind of
    stream."""
    import threading

    import numpy as np

    import particle_fluids as pf
    from particle_fluids import scenes

    dt = 1.0 / 60.0
    rng = np.random.default_rng(args.serve_seed)
    counts = scenes.log_uniform_counts(args.serve_scenes, 5000, 500000, args.serve_seed)
    bodies = rng.integers(0, 31, args.serve_scenes)
    t0 = time.perf_counter()
    srv = pf.Server(context=ctx, hbmBudget=int(args.serve_budget_gb * (1 << 30)), slots=args.serve_slots)
    built, sizes = [], []
    for i, (c, b) in enumerate(zip(counts, bodies)):
        sc = scenes.build_mixed_scene(int(c), int(b), seed=args.serve_seed + i, substeps=args.substeps,
                                      iterations=args.iterations, context=ctx, evictable=True)
        if tile_union:
            sc["fluid"].setTileUnionLimit(tile_union)
        sc["id"] = srv.addScene(sc["loop"])
        built.append(sc)
        sizes.append(sc["ps"].capacity)
        if (i + 1) % 100 == 0:
            print(f"[bench] served: {i + 1} scenes built ({time.perf_counter() - t0:.1f} s, "
                  f"{srv.stats()['resident_bytes'] / 2**30:.1f} GiB resident, {device_free_gib():.1f} GiB free, "
                  f"{process_mappings()} memory mappings)", file=sys.stderr, flush=True)
    sizes = np.asarray(sizes, np.int64)
    st0 = srv.stats()
    print(f"[bench] served: {len(built)} scenes, {int(sizes.sum())} particles, {st0['resident_bytes'] / 2**30:.1f} "
          f"GiB resident", file=sys.stderr, flush=True)

    def run_stream(n, rate, seed):
        r = np.random.default_rng(seed)
        targets = r.integers(0, len(built), n)
        arrivals = np.cumsum(r.exponential(1.0 / rate, n)) if rate > 0 else np.zeros(n)
        first = srv.stats()["submitted"]

        def feeder():
            base = time.perf_counter()
            for k in range(n):
                while True:
                    late = arrivals[k] - (time.perf_counter() - base)
                    if late <= 0:
                        break
                    if late > 3e-4:
                        time.sleep(late - 2e-4)
                srv.submit(built[targets[k]]["id"], dt)

        th = threading.Thread(target=feeder)
        th.start()
        th.join()
        srv.drain()
        info = srv.requests(first, n)
        lat = (info["t_done"] - info["t_submit"]) * 1e3
        span = info["t_done"].max() - info["t_submit"].min()
        work = float(sizes[targets].sum()) * args.substeps
        return {"requests": n, "offered_req_per_s": rate if rate > 0 else None,
                "particle_substeps_per_s": round(work / span, 1), "scene_steps_per_s": round(n / span, 2),
                "p50_ms": round(float(np.percentile(lat, 50)), 3), "p99_ms": round(float(np.percentile(lat, 99)), 3),
                "mean_launch_scenes": round(float(info["launch_scenes"].mean()), 2), "seconds": round(span, 3),
                "failed": int((info["status"] != 2).sum())}

    run_stream(min(200, 2 * len(built)), 0, args.serve_seed + 1)  # warm-up
    poisson = run_stream(args.serve_requests, args.serve_rate, args.serve_seed + 2)
    print(f"[bench] served: Poisson {poisson}", file=sys.stderr, flush=True)
    saturated = run_stream(args.serve_requests, 0, args.serve_seed + 3)
    print(f"[bench] served: saturated {saturated}", file=sys.stderr, flush=True)
    st = srv.stats()
    for sc in built:
        srv.removeScene(sc["id"])
    srv.dispose()

    # The same scenes under a budget below their resident size: admission
    # evicts to host memory, and the stream restores and evicts as it goes.
    evicting = None
    if args.serve_evict_frac > 0:
        budget = int(st0["resident_bytes"] * args.serve_evict_frac)
        srv = pf.Server(context=ctx, hbmBudget=budget, slots=args.serve_slots)
        for sc in built:
            sc["id"] = srv.addScene(sc["loop"])
        ste = srv.stats()
        print(f"[bench] served, evicting: budget {budget / 2**30:.1f} GiB, {ste['evictions']} evictions at "
              f"admission, {ste['host_bytes'] / 2**30:.1f} GiB in host memory", file=sys.stderr, flush=True)
        # capacity first (every request queued at once), then a Poisson stream
        # offered at half of it, so the latencies are not a queue's
        saturated_e = run_stream(args.serve_evict_requests, 0, args.serve_seed + 5)
        print(f"[bench] served, evicting: saturated {saturated_e}", file=sys.stderr, flush=True)
        rate_e = max(1.0, round(0.5 * saturated_e["scene_steps_per_s"], 1))
        poisson_e = run_stream(args.serve_evict_requests, rate_e, args.serve_seed + 6)
        st1 = srv.stats()
        evicting = {"hbm_budget_gib": round(budget / 2**30, 2), "saturated": saturated_e, "poisson": poisson_e,
                    "evictions": st1["evictions"], "restores": st1["restores"],
                    "evict_failures": st1["evict_failures"], "host_gib": round(st1["host_bytes"] / 2**30, 2),
                    "peak_resident_gib": round(st1["peak_resident_bytes"] / 2**30, 2)}
        print(f"[bench] served, evicting: {evicting}", file=sys.stderr, flush=True)
        for sc in built:
            srv.removeScene(sc["id"])
        srv.dispose()

    # plain form: the same kind of stream, one loop step per request
    r = np.random.default_rng(args.serve_seed + 4)
    targets = r.integers(0, len(built), args.serve_plain_requests)
    for sc in built:  # capture every loop's graph before the timed stream
        sc["loop"].useGraph(True)
        sc["loop"].step(dt)
    ctx.synchronize()
    tp = time.perf_counter()
    for k in targets:
        built[k]["loop"].step(dt)
    ctx.synchronize()
    tp = time.perf_counter() - tp
    plain = {"requests": int(len(targets)),
             "particle_substeps_per_s": round(float(sizes[targets].sum()) * args.substeps / tp, 1),
             "scene_steps_per_s": round(len(targets) / tp, 2)}
    print(f"[bench] served: plain {plain}", file=sys.stderr, flush=True)
    for sc in built:
        scenes.dispose_scene(sc)
    return {
        "workload": f"{len(built)} scenes, fluid budgets log-uniform in [5000, 500000], 0-30 soft cubes of 2197 "
                    f"particles with particle contacts (seed {args.serve_seed}), {args.substeps} substeps x "
                    f"{args.iterations} iterations, "
                    f"step(1/60) requests on uniformly chosen scenes",
        "scenes": len(built), "particles": int(sizes.sum()),
        "hbm_budget_gib": args.serve_budget_gb, "resident_gib": round(st0["resident_bytes"] / 2**30, 2),
        "slots": args.serve_slots, "tile_union_capacity": tile_union, "poisson": poisson, "saturated": saturated, "plain": plain,
        "speedup_saturated_vs_plain": round(saturated["particle_substeps_per_s"] / plain["particle_substeps_per_s"], 3),
        "launches": st["launches"], "evictions": st["evictions"], "restores": st["restores"],
        "evicting": evicting,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--budget", type=int, default=1 << 20)
    ap.add_argument("--substeps", type=int, default=2)
    ap.add_argument("--iterations", type=int, default=4)
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--no-autotune", action="store_true", help="batched form: keep the default tile-union capacity")
    ap.add_argument("--autotune-plain", action="store_true",
                    help="tune the single scene's tile-union capacity on a twin (the default 2048 won its sweep)")
    ap.add_argument("--cpu-baseline-steps", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-c1-steps", type=int, default=200, help="C1 dam break steps on the CPU oracle (0: skip)")
    ap.add_argument("--sort-reps", type=int, default=20, help="isolated 4M sort repetitions timed (0: skip)")
    ap.add_argument("--batch-scenes", type=int, default=256)
    ap.add_argument("--batch-steps", type=int, default=10)
    ap.add_argument("--batch-warmup", type=int, default=3)
    ap.add_argument("--batch-seed", type=int, default=0xBA7C4)
    ap.add_argument("--serve-scenes", type=int, default=1000)
    ap.add_argument("--serve-requests", type=int, default=4000)
    ap.add_argument("--serve-plain-requests", type=int, default=1000)
    ap.add_argument("--serve-rate", type=float, default=2000.0)
    ap.add_argument("--serve-budget-gb", type=float, default=200.0)
    ap.add_argument("--serve-slots", type=int, default=2)
    ap.add_argument("--serve-seed", type=int, default=0x5E7E)
    ap.add_argument("--serve-evict-requests", type=int, default=400,
                    help="requests per stream of the evicting phase (saturated, then Poisson at half its rate)")
    ap.add_argument("--serve-evict-frac", type=float, default=0.6,
                    help="then serve the same scenes under this fraction of their resident bytes (0: skip)")
    ap.add_argument("--field-updates", type=int, default=20, help="surface-field updates timed (0: skip)")
    ap.add_argument("--viscosity-steps", type=int, default=20, help="thick-liquid steps timed (0: skip)")
    ap.add_argument("--viscosity-sweeps", type=int, default=12)
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="end with a device copy of this size (a known-traffic kernel for HBM counters)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus above 1 is not supported (scenes shard "
                                   "across GPUs without exchange; one scene per GPU)"}))
        sys.exit(2)

    import particle_fluids as pf
    from particle_fluids import scenes

    ctx = pf.default_context()
    dt = 1.0 / 60.0
    print(f"[bench] building dam break, budget {args.budget}", file=sys.stderr, flush=True)
    sc = scenes.build_dam_break(args.budget, substeps=args.substeps, iterations=args.iterations,
                                use_graph=not args.no_graph)
    loop = sc["loop"]
    n = sc["ps"].capacity
    H = loop.info()["hashTableSize"]
    tuned = None
    if args.autotune_plain:
        # tile-union LDS capacity, tuned on a twin of the scene at the same
        # point of its run, so the measured scene keeps its configured window
        twin = scenes.build_dam_break(args.budget, substeps=args.substeps, iterations=args.iterations, use_graph=True)
        for _ in range(args.warmup):
            twin["loop"].step(dt)
        best, times = twin["loop"].autotune(dt, steps=2)
        scenes.dispose_scene(twin)
        sc["fluid"].setTileUnionLimit(best)
        tuned = {"tile_union_capacity": best, "ms_per_step": {str(k): round(v, 3) for k, v in times.items()}}
        print(f"[bench] autotune: {tuned}", file=sys.stderr, flush=True)
    for _ in range(args.warmup):
        loop.step(dt)
    ctx.synchronize()
    print(f"[bench] N={n} H={H}, warmup {args.warmup} done, timing {args.steps} steps", file=sys.stderr, flush=True)

    ctx.timer_begin()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        loop.step(dt)
    gpu_ms = ctx.timer_end()
    wall = time.perf_counter() - t0
    seconds = max(wall, gpu_ms / 1000.0)
    substeps_done = args.steps * args.substeps
    value = n * substeps_done / seconds
    ov = loop.readbackOverflow()
    print(f"[bench] {seconds:.3f} s, {value:.4e} particle-substeps/s, overflow {ov}", file=sys.stderr, flush=True)

    # Per-kernel durations of the timed window: a twin of the scene, stepped
    # through the same warm-up with its graph, then launched kernel by kernel
    # with HIP events around every launch over the same steps.
    twin = scenes.build_dam_break(args.budget, substeps=args.substeps, iterations=args.iterations,
                                  use_graph=not args.no_graph)
    for _ in range(args.warmup):
        twin["loop"].step(dt)
    twin["loop"].useGraph(False)
    ctx.synchronize()
    ctx.set_profiling(True)
    ctx.profile_reset()
    for _ in range(args.steps):
        twin["loop"].step(dt)
    stats = ctx.profile()
    ctx.set_profiling(False)
    ctx.profile_reset()
    twin_ov = twin["loop"].readbackOverflow()
    scenes.dispose_scene(twin)
    profiled_ms = sum(t for _, t in stats.values())
    dominant = max(stats, key=lambda k: stats[k][1])
    launches, total = stats[dominant]
    avg_s = total / launches / 1000.0
    bpp = KERNEL_BYTES.get(dominant)
    kernels = {k: {"launches": v[0], "avg_us": round(v[1] / v[0] * 1000.0, 3),
                   "share": round(v[1] / profiled_ms, 4)} for k, v in sorted(stats.items(), key=lambda kv: -kv[1][1])}
    roofline = None
    if bpp is not None:
        achieved = bpp * n / avg_s / 1e9
        tb = HBM_TRAFFIC_B_PER_PARTICLE.get(dominant)
        roofline = {"bound": "hbm", "kernel": dominant, "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                    "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5),
                    "traffic": round(tb * n) if tb is not None and args.budget == 1 << 20 else None,
                    "bytes_per_launch": bpp * n, "avg_launch_us": round(avg_s * 1e6, 3)}
    roofline_valu = None
    vpp = VALU_INSTS_PER_PARTICLE.get(dominant)
    if vpp is not None and args.budget == 1 << 20:
        rate = vpp * n / avg_s
        roofline_valu = {"bound": "valu", "kernel": dominant, "achieved": round(rate / 1e9, 2),
                         "peak": round(VALU_PEAK_WAVE_INSTS / 1e9, 1), "unit": "G wave-instructions/s",
                         "frac": round(rate / VALU_PEAK_WAVE_INSTS, 4),
                         "instructions_per_launch": round(vpp * n)}
    sub_bytes = SUBSTEP_BYTES_PER_PARTICLE(args.iterations) * n + SCAN_BYTES_PER_BUCKET * H
    sub_s = seconds / substeps_done
    roofline_substep = {"bound": "hbm", "achieved": round(sub_bytes / sub_s / 1e9, 2), "peak": HBM_PEAK_GBS,
                        "unit": "GB/s", "frac": round(sub_bytes / sub_s / 1e9 / HBM_PEAK_GBS, 5),
                        "bytes_per_substep": sub_bytes}
    roofline_valu_substep = None
    if args.budget == 1 << 20:
        insts = VALU_SUBSTEP_INSTS_PER_PARTICLE(args.iterations) * n
        roofline_valu_substep = {"bound": "valu", "achieved": round(insts / sub_s / 1e9, 2),
                                 "peak": round(VALU_PEAK_WAVE_INSTS / 1e9, 1), "unit": "G wave-instructions/s",
                                 "frac": round(insts / sub_s / VALU_PEAK_WAVE_INSTS, 4),
                                 "instructions_per_substep": round(insts)}

    cpu = None
    if not args.no_cpu_baseline:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as orc

        # The same scene at the end of the timed window (positions and
        # velocities read back), so the CPU steps cost what the GPU's did.
        snap = sc["ps"].readback()
        threads = host_threads()
        orc.set_threads(threads)
        o = oracle_dam_break(orc, snap.positions[:, :3], snap.velocities[:, :3], sc["radius"], args.substeps,
                             args.iterations, H)
        print(f"[bench] CPU baseline: {args.cpu_baseline_steps} step(s) of the oracle on {threads} threads",
              file=sys.stderr, flush=True)
        c0 = time.perf_counter()
        for _ in range(args.cpu_baseline_steps):
            o.step(dt)
        csec = time.perf_counter() - c0
        orc.set_threads(0)
        del o
        cpu = {"value": round(n * args.substeps * args.cpu_baseline_steps / csec, 1), "unit": "particle-substeps/s",
               "cores": threads, "kind": "port",
               "sample": f"{args.cpu_baseline_steps} step(s) = {args.substeps * args.cpu_baseline_steps} substeps of the "
                         f"same {n}-particle scene from its state after the timed window (step "
                         f"{args.warmup + args.steps}), FP32 C oracle (OpenMP particle loops), {threads} threads, "
                         f"{csec:.1f} s on {cpu_model()}"}
    cpu_c1 = run_cpu_c1(args) if args.cpu_c1_steps > 0 else None
    sort = run_sort(args, ctx) if args.sort_reps > 0 else None

    field = run_surface_field(args, ctx, sc) if args.field_updates > 0 else None
    viscosity = run_viscosity(args, ctx, dt) if args.viscosity_steps > 0 else None

    batched = None
    if args.batch_scenes > 0:
        batched = run_batched(args, ctx)
    served = None
    if args.serve_scenes > 0:
        served = run_served(args, ctx)

    line = {
        "metric": "particle-substeps/s at 1M PBF",
        "value": round(value, 1),
        "unit": "particle-substeps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(seconds / args.steps * 1000.0, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"dam break, plain method, {n} particles (budget {args.budget}), "
                               f"{args.substeps} substeps x {args.iterations} iterations, dt 1/60, one scene",
                   "particles": n, "hash_table_size": H, "substeps": args.substeps,
                   "iterations": args.iterations, "graph": not args.no_graph, "autotune": tuned},
        "roofline": roofline,
        "roofline_valu": roofline_valu,
        "roofline_valu_substep": roofline_valu_substep,
        "roofline_substep": roofline_substep,
        "overflow": {"timed_window": dict(ov._asdict()), "profiled_twin": dict(twin_ov._asdict())},
        "cpu_baseline": cpu,
        "cpu_c1": cpu_c1,
        "sort": sort,
        "kernels": kernels,
        "gpu_ms_timed": round(gpu_ms, 3),
        "surface_field": field,
        "viscosity": viscosity,
        "batched": batched,
        "served": served,
    }
    print(json.dumps(line), flush=True)
    if args.copy_calibration_mib:
        calibration_copy(args.copy_calibration_mib << 20)


def calibration_copy(nbytes):
    """A fill and a device-to-device copy of known size through the HIP
    runtime (its fill and copy kernels): their FETCH_SIZE / WRITE_SIZE
    readings calibrate the HBM counters against the bytes they must move."""
    import ctypes

    hip = ctypes.CDLL("libamdhip64.so")
    vp = ctypes.c_void_p
    hip.hipMalloc.argtypes = [ctypes.POINTER(vp), ctypes.c_size_t]
    hip.hipMemset.argtypes = [vp, ctypes.c_int, ctypes.c_size_t]
    hip.hipMemcpy.argtypes = [vp, vp, ctypes.c_size_t, ctypes.c_int]
    hip.hipFree.argtypes = [vp]
    src, dst = vp(), vp()
    for p in (src, dst):
        if hip.hipMalloc(ctypes.byref(p), nbytes) != 0:
            raise RuntimeError("calibration: hipMalloc failed")
    rc = hip.hipMemset(src, 1, nbytes) or hip.hipMemcpy(dst, src, nbytes, 3) or hip.hipDeviceSynchronize()
    hip.hipFree(src)
    hip.hipFree(dst)
    if rc:
        raise RuntimeError(f"calibration: HIP error {rc}")
    print(f"[bench] calibration: fill of {nbytes} bytes, then copy of {nbytes} bytes", file=sys.stderr, flush=True)


if __name__ == "__main__":
    main()
