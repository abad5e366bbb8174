"""The benchmark configurations at their configured sizes, on the GPU.

- The 1M dam break (budget 2^20: 1,028,772 particles, 2^20 buckets) against
  the FP32 oracle, bit for bit.
- 256 batched scenes of 5,000-200,000 particles (every other one with 20
  soft cubes and contacts): a sample spanning the sizes equals the same
  scenes stepped alone, bit for bit.
- 1,000 served scenes of 5,000-500,000 particles with 0-30 cubes, under an
  HBM budget below their resident size so scenes are evicted to host memory
  and restored: a sample equals the same scenes stepped alone.
The 4,194,304-point isolated sort is checked in test_gpu_grid.py.
"""
import numpy as np
import pytest

import oracle as orc
import particle_fluids as pf
from helpers import oracle_dam_break
from particle_fluids import scenes

pytestmark = pytest.mark.gpu

DT = 1.0 / 60.0


def state(sc):
    s = sc["ps"].readback()
    out = dict(x=s.positions, v=s.velocities)
    if sc.get("soft") is not None:
        c, r = sc["soft"].readbackFrames()
        out.update(c=c, r=r)
    return out


def assert_same(a, b, what):
    for key in a:
        assert np.array_equal(a[key], b[key]), (what, key)


def spanning_sample(counts, k):
    """k scene indices spread over the size range (smallest to largest)."""
    order = np.argsort(counts, kind="stable")
    return sorted(set(int(order[int(round(q))]) for q in np.linspace(0, len(order) - 1, k)))


def test_c2_dam_break_1m_matches_fp32_oracle():
    print("[configs] C2: 1M dam break, 2 steps vs the FP32 oracle", flush=True)
    sc = scenes.build_dam_break(1 << 20, substeps=2, iterations=4, use_graph=True)
    n = sc["ps"].capacity
    H = sc["loop"].info()["hashTableSize"]
    assert n == 1028772 and H == 1 << 20
    orc.set_threads(0)
    o = oracle_dam_break(sc["positions"], sc["radius"], 2, 4, "f32", H=H)
    for step in range(2):
        sc["loop"].step(DT)
        o.step(DT)
        snap = sc["ps"].readback()
        ref = o.get()
        rho = sc["fluid"].readbackDensity()
        same = {"x": np.array_equal(snap.positions[:, :3].astype(np.float64), ref["x"]),
                "v": np.array_equal(snap.velocities[:, :3].astype(np.float64), ref["v"]),
                "rho": np.array_equal(rho.astype(np.float64), ref["rho"])}
        print(f"[configs] C2 step {step + 1}: bit-identical {same}", flush=True)
        assert all(same.values()), same
    scenes.dispose_scene(sc)


def test_c4_batch_of_256_scenes_matches_scenes_alone():
    seed = 0xBA7C4
    counts = scenes.log_uniform_counts(256, 5000, 200000, seed)
    print(f"[configs] C4: 256 scenes, {int(counts.sum())} fluid budget", flush=True)
    built = [scenes.build_mixed_scene(int(c), 20 if i % 2 else 0, seed=seed + i) for i, c in enumerate(counts)]
    sample = spanning_sample(counts, 16)
    sample += [i for i in (1, 2) if i not in sample]  # both kinds for sure
    alone = {i: scenes.build_mixed_scene(int(counts[i]), 20 if i % 2 else 0, seed=seed + i) for i in sample}
    batch = pf.Batch([b["loop"] for b in built])
    batch.useGraph(True)
    for _ in range(2):
        batch.step(DT)
        for a in alone.values():
            a["loop"].step(DT)
    print(f"[configs] C4: comparing scenes {sample}", flush=True)
    for i in sample:
        assert_same(state(built[i]), state(alone[i]), i)
    batch.dispose()
    for b in built + list(alone.values()):
        scenes.dispose_scene(b)


def test_c5_thousand_served_scenes_with_evictions_match_alone():
    seed = 0x5E7E
    counts = scenes.log_uniform_counts(1000, 5000, 500000, seed)
    bodies = np.random.default_rng(seed).integers(0, 31, 1000)
    served = []
    for i, (c, b) in enumerate(zip(counts, bodies)):
        served.append(scenes.build_mixed_scene(int(c), int(b), seed=seed + i, evictable=True))
    resident = sum(s["ps"].deviceBytes() for s in served)
    budget = int(resident * 0.6)
    print(f"[configs] C5: 1000 scenes, {sum(s['ps'].capacity for s in served)} particles, "
          f"{resident / 2**30:.1f} GiB resident, budget {budget / 2**30:.1f} GiB", flush=True)
    srv = pf.Server(hbmBudget=budget, slots=2)
    ids = [srv.addScene(s["loop"]) for s in served]
    st0 = srv.stats()
    assert st0["evictions"] > 0 and st0["resident_bytes"] <= budget
    sample = spanning_sample(counts, 16)
    rng = np.random.default_rng(seed + 1)
    targets = np.concatenate([rng.integers(0, len(ids), 1500), np.repeat(sample, 2)])
    rng.shuffle(targets)
    for t in targets:
        srv.submit(ids[t], DT)
    srv.drain()
    st = srv.stats()
    print(f"[configs] C5 stats {st}", flush=True)
    assert st["failed"] == 0 and st["completed"] == len(targets)
    assert st["evictions"] > st0["evictions"] and st["restores"] > 0
    assert st["peak_resident_bytes"] <= budget
    steps = {i: int((targets == i).sum()) for i in sample}
    for i in sample:
        srv.removeScene(ids[i])
    srv.dispose()
    print(f"[configs] C5: comparing scenes {sample} after {steps} steps", flush=True)
    for i in sample:
        a = scenes.build_mixed_scene(int(counts[i]), int(bodies[i]), seed=seed + i)
        for _ in range(steps[i]):
            a["loop"].step(DT)
        assert_same(state(served[i]), state(a), i)
        scenes.dispose_scene(a)
    for s in served:
        scenes.dispose_scene(s)
