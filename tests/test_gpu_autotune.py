"""The autotuner sets a tile-union LDS capacity; every capacity gives the
same results (the union only changes where particles are staged)."""
import numpy as np
import pytest

import particle_fluids as pf
from particle_fluids import scenes

pytestmark = pytest.mark.gpu

DT = 1.0 / 60.0


@pytest.mark.parametrize("cap", [256, 768, 1280])
def test_capacity_does_not_change_results(cap):
    print(f"[autotune] capacity {cap}", flush=True)
    a = scenes.build_dam_break(20000, seed=3)
    b = scenes.build_dam_break(20000, seed=3, use_graph=True)
    b["fluid"].setTileUnionLimit(cap)
    for _ in range(3):
        a["loop"].step(DT)
        b["loop"].step(DT)
    sa, sb = a["ps"].readback(), b["ps"].readback()
    assert np.array_equal(sa.positions, sb.positions) and np.array_equal(sa.velocities, sb.velocities)


def test_loop_and_batch_autotune():
    sc = scenes.build_dam_break(20000, seed=4, use_graph=True)
    best, times = sc["loop"].autotune(DT, candidates=(512, 1024, 2048), steps=1)
    print(f"[autotune] loop best {best} {times}", flush=True)
    assert best in (512, 1024, 2048) and set(times) == {512, 1024, 2048}
    assert all(t > 0 for t in times.values())
    mixed = [scenes.build_mixed_scene(b, nb, seed=50 + i) for i, (b, nb) in enumerate([(3000, 1), (8000, 0)])]
    alone = [scenes.build_mixed_scene(b, nb, seed=50 + i) for i, (b, nb) in enumerate([(3000, 1), (8000, 0)])]
    batch = pf.Batch([m["loop"] for m in mixed])
    batch.useGraph(True)
    best, times = batch.autotune(DT, candidates=(768, 1536), steps=1)
    print(f"[autotune] batch best {best} {times}", flush=True)
    # 2 sweeps x 2 candidates x (1 capture + 1 timed) steps, then compare with the scenes alone
    for s in alone:
        for _ in range(8):
            s["loop"].step(DT)
    for a, m in zip(alone, mixed):
        ra, rm = a["ps"].readback(), m["ps"].readback()
        assert np.array_equal(ra.positions, rm.positions)
    with pytest.raises(pf.PFError, match="outside"):
        batch.autotune(DT, candidates=(100,))
