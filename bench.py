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
`roofline_substep` applies the same byte model to the whole substep.
`cpu_baseline` times the CPU oracle (the FP32 restatement of the reference
schedule, one thread) on one step of the same scene.
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

# Algorithmic bytes per particle per launch of each kernel, by the byte model
# in DESIGN.md (FP32 vec3 = 12 B, scalar = 4 B; neighbour gathers, neighbour
# lists and collider multipliers count 0; per-bucket arrays 4 B a touch).
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
SCAN_BYTES_PER_BUCKET = 12


def run_batched(args, ctx):
    """256 independent scenes, particle budgets log-uniform in [5,000, 200,000];
    every other scene adds 20 soft cubes of 2,197 particles (global shape
    matching, stiffness uniform in [0.1, 1.0]). Times the same scenes stepped
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
        for L in loops:
            L.step(dt)
        batch.step(dt)
    ctx.synchronize()

    def timed(fn):
        ctx.synchronize()
        t = time.perf_counter()
        for _ in range(args.batch_steps):
            fn()
        ctx.synchronize()
        return time.perf_counter() - t

    def plain_step():
        for L in loops:
            L.step(dt)

    t_plain = timed(plain_step)
    t_batch = timed(lambda: batch.step(dt))
    t_plain2 = timed(plain_step)
    t_plain = 0.5 * (t_plain + t_plain2)
    work = n_total * args.substeps * args.batch_steps
    out = {
        "workload": f"{len(built)} scenes, budgets log-uniform in [5000, 200000] (seed {args.batch_seed}), "
                    f"every other scene + 20 soft cubes of 2197 particles, {args.substeps} substeps x "
                    f"{args.iterations} iterations, dt 1/60",
        "scenes": len(built), "particles": n_total, "fluid_particles": n_fluid,
        "steps": args.batch_steps,
        "plain": {"particle_substeps_per_s": round(work / t_plain, 1),
                  "scene_steps_per_s": round(len(built) * args.batch_steps / t_plain, 2),
                  "ms_per_round": round(t_plain / args.batch_steps * 1e3, 3)},
        "batched": {"particle_substeps_per_s": round(work / t_batch, 1),
                    "scene_steps_per_s": round(len(built) * args.batch_steps / t_batch, 2),
                    "ms_per_step": round(t_batch / args.batch_steps * 1e3, 3)},
        "speedup": round(t_plain / t_batch, 3),
    }
    print(f"[bench] batched: plain {out['plain']}, batched {out['batched']}, x{out['speedup']}",
          file=sys.stderr, flush=True)
    batch.dispose()
    for b in built:
        scenes.dispose_scene(b)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--budget", type=int, default=1 << 20)
    ap.add_argument("--substeps", type=int, default=2)
    ap.add_argument("--iterations", type=int, default=4)
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--profile-steps", type=int, default=10)
    ap.add_argument("--cpu-baseline-steps", type=int, default=1)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--batch-scenes", type=int, default=256)
    ap.add_argument("--batch-steps", type=int, default=10)
    ap.add_argument("--batch-warmup", type=int, default=3)
    ap.add_argument("--batch-seed", type=int, default=0xBA7C4)
    args = ap.parse_args()

    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus above 1 is not supported (scenes shard "
                                   "across GPUs without exchange, see DESIGN.md)"}))
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

    # Per-kernel durations: the same workload launched kernel by kernel with
    # HIP events around every launch on the library's stream.
    loop.useGraph(False)
    ctx.set_profiling(True)
    ctx.profile_reset()
    for _ in range(args.profile_steps):
        loop.step(dt)
    stats = ctx.profile()
    ctx.set_profiling(False)
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
        roofline = {"bound": "hbm", "kernel": dominant, "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                    "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": None,
                    "bytes_per_launch": bpp * n, "avg_launch_us": round(avg_s * 1e6, 3)}
    sub_bytes = SUBSTEP_BYTES_PER_PARTICLE(args.iterations) * n + SCAN_BYTES_PER_BUCKET * H
    sub_s = seconds / substeps_done
    roofline_substep = {"bound": "hbm", "achieved": round(sub_bytes / sub_s / 1e9, 2), "peak": HBM_PEAK_GBS,
                        "unit": "GB/s", "frac": round(sub_bytes / sub_s / 1e9 / HBM_PEAK_GBS, 5),
                        "bytes_per_substep": sub_bytes}

    cpu = None
    if not args.no_cpu_baseline:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as orc

        pos, r = sc["positions"], sc["radius"]
        o = orc.Scene(n, r, "f32")
        o.set_particles(pos)
        o.set_loop(args.substeps, args.iterations, (0.0, -9.81, 0.0), 4 * r, H)
        o.add_fluid(0, n, 1000.0, 2 * r, 4 * r, 1e-4, viscosity=0.01, vorticity=0.02, surface_tension=0.1,
                    adhesion=0.1)
        mn, mx = scenes.CONTAINER[:3], scenes.CONTAINER[3:]
        for nrm, which in scenes.PLANES:
            o.add_plane(nrm, mn if which == "min" else mx)
        print(f"[bench] CPU baseline: {args.cpu_baseline_steps} step(s) of the oracle", file=sys.stderr, flush=True)
        c0 = time.perf_counter()
        for _ in range(args.cpu_baseline_steps):
            o.step(dt)
        csec = time.perf_counter() - c0
        cpu = {"value": round(n * args.substeps * args.cpu_baseline_steps / csec, 1), "unit": "particle-substeps/s",
               "cores": 1, "kind": "port",
               "sample": f"{args.cpu_baseline_steps} step(s) = {args.substeps * args.cpu_baseline_steps} substeps of the "
                         f"same {n}-particle scene, FP32 C oracle, 1 thread, {csec:.1f} s on {platform.processor() or 'host'}"}

    batched = None
    if args.batch_scenes > 0:
        batched = run_batched(args, ctx)

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
                   "iterations": args.iterations, "graph": not args.no_graph},
        "roofline": roofline,
        "roofline_substep": roofline_substep,
        "cpu_baseline": cpu,
        "kernels": kernels,
        "gpu_ms_timed": round(gpu_ms, 3),
        "batched": batched,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
