"""Scenes of the batched and served mixes (dam break + globally shape-matched
cubes in an extended tank): one scene against the FP32 oracle, and a mixed
batch against each scene run alone."""
import numpy as np
import pytest

import particle_fluids as pf
from particle_fluids import scenes
from helpers import oracle_mixed_scene, rel_err

pytestmark = pytest.mark.gpu

TOL = 1e-5


def state(sc):
    s = sc["ps"].readback()
    return dict(x=s.positions, v=s.velocities)


def test_mixed_scene_matches_fp32_oracle():
    print("[batched] mixed scene vs oracle: budget 3000, 2 cubes", flush=True)
    sc = scenes.build_mixed_scene(3000, 2, seed=11)
    H = sc["loop"].info()["hashTableSize"]
    o = oracle_mixed_scene(sc["layout"], 2, 4, H)
    errs = []
    for k in range(3):
        sc["loop"].step(1.0 / 60.0)
        o.step(1.0 / 60.0)
        g, ref = state(sc), o.get()
        errs.append((rel_err(g["x"][:, :3], ref["x"]), rel_err(g["v"][:, :3], ref["v"])))
    print(f"[batched] rel err per step (x, v) {errs}", flush=True)
    for ex, ev in errs:
        assert ex <= TOL and ev <= 1e-4, errs


@pytest.mark.parametrize("graph", [False, True])
def test_mixed_batch_matches_each_scene_alone(graph):
    # fluid-only scenes and soft scenes with contacts in one batch (each
    # schedule a group of the batch): the results must still be bit-identical
    spec = [(3000, 0), (6000, 2), (4000, 3), (9000, 0), (2500, 1)]
    print(f"[batched] mixed batch vs alone, graph={graph}, {spec}", flush=True)
    alone = [scenes.build_mixed_scene(b, nb, seed=100 + i) for i, (b, nb) in enumerate(spec)]
    mixed = [scenes.build_mixed_scene(b, nb, seed=100 + i) for i, (b, nb) in enumerate(spec)]
    batch = pf.Batch([s["loop"] for s in mixed])
    batch.useGraph(graph)
    for _ in range(3):
        for s in alone:
            s["loop"].step(1.0 / 60.0)
        batch.step(1.0 / 60.0)
    for i, (s, t) in enumerate(zip(alone, mixed)):
        ga, gb = state(s), state(t)
        for key in ga:
            assert np.array_equal(ga[key], gb[key]), (i, key)
        if s["soft"] is not None:
            ca, ra = s["soft"].readbackFrames()
            cb, rb = t["soft"].readbackFrames()
            assert np.array_equal(ca, cb) and np.array_equal(ra, rb), i


@pytest.mark.parametrize("kinds", [3, 6])
def test_batch_of_different_schedules_matches_each_scene_alone(kinds):
    # scenes whose fluids carry different effects, beside scenes with cubes:
    # each schedule runs as its own group of the batch (3 kinds), or, past
    # four different schedules, all under the merged one (6 kinds)
    options = [dict(scenes.FLUID_DEFAULTS), dict(viscosity=0.01, vorticity=0.02),
               dict(surfaceTension=0.1), dict(viscosity=0.01), dict(vorticity=0.02, adhesion=0.1), {}]
    spec = [(3000 + 500 * i, 2 if i == 1 else 0, options[i if i != 1 else 0]) for i in range(kinds)]
    spec += [(2500, 0, options[0]), (3500, 1, options[0])]
    print(f"[batched] {kinds} schedules vs alone: {[(b, nb, sorted(o)) for b, nb, o in spec]}", flush=True)
    build = lambda i, b, nb, o: scenes.build_mixed_scene(b, nb, seed=500 + i, fluid_options=o)  # noqa: E731
    alone = [build(i, *x) for i, x in enumerate(spec)]
    mixed = [build(i, *x) for i, x in enumerate(spec)]
    batch = pf.Batch([s["loop"] for s in mixed])
    batch.useGraph(True)
    for _ in range(3):
        for s in alone:
            s["loop"].step(1.0 / 60.0)
        batch.step(1.0 / 60.0)
    for i, (s, t) in enumerate(zip(alone, mixed)):
        ga, gb = state(s), state(t)
        for key in ga:
            assert np.array_equal(ga[key], gb[key]), (i, key)
    batch.dispose()
