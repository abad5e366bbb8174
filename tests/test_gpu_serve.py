"""The served form: scenes in arenas of reserved device addresses, a request
loop with continuous batching, HBM-budget admission and eviction to host.
Every scene's result must be the same alone and in the mix (bit-identical:
the launch covers the same kernels over the same buffers)."""
import numpy as np
import pytest

import particle_fluids as pf
from particle_fluids import scenes

pytestmark = pytest.mark.gpu

DT = 1.0 / 60.0
SPEC = [(3000, 0), (6000, 2), (4000, 1), (9000, 0), (2500, 1), (5000, 0)]


def state(sc):
    s = sc["ps"].readback()
    out = dict(x=s.positions, v=s.velocities)
    if sc["soft"] is not None:
        c, r = sc["soft"].readbackFrames()
        out.update(c=c, r=r)
    return out


def assert_same(a, b, what):
    for key in a:
        assert np.array_equal(a[key], b[key]), (what, key)


def twins(spec, seed0=200):
    alone = [scenes.build_mixed_scene(b, nb, seed=seed0 + i) for i, (b, nb) in enumerate(spec)]
    served = [scenes.build_mixed_scene(b, nb, seed=seed0 + i, evictable=True) for i, (b, nb) in enumerate(spec)]
    return alone, served


def test_evictable_scene_steps_like_a_plain_one():
    print("[serve] evictable vs plain scene", flush=True)
    a = scenes.build_mixed_scene(4000, 1, seed=5)
    b = scenes.build_mixed_scene(4000, 1, seed=5, evictable=True)
    assert b["ps"].deviceBytes() > 0 and a["ps"].deviceBytes() == 0
    for _ in range(3):
        a["loop"].step(DT)
        b["loop"].step(DT)
    assert_same(state(a), state(b), "evictable")


def test_served_scenes_match_alone_and_batch():
    print(f"[serve] served mix vs alone {SPEC}", flush=True)
    alone, served = twins(SPEC)
    srv = pf.Server(slots=2)
    ids = [srv.addScene(s["loop"]) for s in served]
    rng = np.random.default_rng(1)
    tickets = []
    for _ in range(3):
        for i in rng.permutation(len(ids)):
            tickets.append(srv.submit(ids[i], DT))
    srv.drain()
    st = srv.stats()
    info = srv.requests(0, len(tickets))
    print(f"[serve] stats {st}", flush=True)
    assert st["completed"] == 3 * len(ids) and st["failed"] == 0
    assert st["max_launch_scenes"] > 1, "requests queued together must share a launch"
    assert np.all(info["status"] == 2)
    assert np.all(info["t_submit"] <= info["t_launch"]) and np.all(info["t_launch"] <= info["t_done"])
    with pytest.raises(pf.PFError, match="being served"):
        served[0]["loop"].step(DT)
    with pytest.raises(pf.PFError, match="being served"):
        served[0]["ps"].readback()
    for i in ids:
        srv.removeScene(i)
    for s in alone:
        for _ in range(3):
            s["loop"].step(DT)
    for k, (a, b) in enumerate(zip(alone, served)):
        assert_same(state(a), state(b), k)
    # handed back: the owner steps it again as a plain loop
    alone[0]["loop"].step(DT)
    served[0]["loop"].step(DT)
    assert_same(state(alone[0]), state(served[0]), "after removal")
    srv.dispose()


def test_small_budget_evicts_and_restores_exactly():
    spec = SPEC[:5]
    alone, served = twins(spec, seed0=300)
    sizes = [s["ps"].deviceBytes() for s in served]
    budget = sorted(sizes)[-1] + sorted(sizes)[-2]
    print(f"[serve] budget {budget} for scene bytes {sizes}", flush=True)
    srv = pf.Server(hbmBudget=budget, slots=2)
    ids = [srv.addScene(s["loop"]) for s in served]
    st0 = srv.stats()
    assert st0["resident_bytes"] <= budget and st0["evictions"] >= 1
    assert st0["resident_scenes"] < len(spec)
    for _ in range(3):
        for i in ids:
            srv.submit(i, DT)
    srv.drain()
    srv.evictScene(ids[0])
    srv.submit(ids[0], DT)
    srv.drain()
    st = srv.stats()
    print(f"[serve] stats {st}", flush=True)
    assert st["failed"] == 0 and st["evictions"] > st0["evictions"] and st["restores"] > 0
    assert st["peak_resident_bytes"] <= budget
    for i in ids:
        srv.removeScene(i)
    for k, s in enumerate(alone):
        for _ in range(4 if k == 0 else 3):
            s["loop"].step(DT)
    for k, (a, b) in enumerate(zip(alone, served)):
        assert_same(state(a), state(b), k)
    srv.dispose()


def test_scene_larger_than_budget_fails_its_request():
    sc = scenes.build_mixed_scene(3000, 0, seed=9, evictable=True)
    plain = scenes.build_mixed_scene(3000, 0, seed=9)
    srv = pf.Server(hbmBudget=1 << 20)
    sid = srv.addScene(sc["loop"])
    assert srv.stats()["resident_scenes"] == 0  # admitted to host memory
    t = srv.submit(sid, DT)
    with pytest.raises(pf.PFError, match="HBM budget"):
        srv.wait(t)
    with pytest.raises(pf.PFError, match="evictable"):
        srv.addScene(plain["loop"])
    srv.removeScene(sid)
    sc["loop"].step(DT)
    plain["loop"].step(DT)
    assert_same(state(sc), state(plain), "after failed request")
    srv.dispose()


def test_refused_eviction_keeps_scenes_and_serving():
    # Host memory for evicted scenes is capped (hostBudget): an eviction past
    # it is refused, the scene stays resident and intact, the request that
    # needed the room fails alone, and serving goes on.
    spec = [(3000, 0), (3000, 0), (3000, 0)]
    alone, served = twins(spec, seed0=400)
    sizes = [s["ps"].deviceBytes() for s in served]
    budget = 2 * max(sizes) + max(sizes) // 2
    # the host bytes one evicted scene takes, from a server without a host cap
    probe = pf.Server(hbmBudget=budget, slots=2)
    pids = [probe.addScene(s["loop"]) for s in served]
    one = probe.stats()["host_bytes"]
    for i in pids:
        probe.removeScene(i)
    probe.dispose()
    host = one + one // 2  # room in host memory for one evicted scene
    print(f"[serve] refused eviction: budget {budget}, host budget {host}, scene bytes {sizes}", flush=True)
    srv = pf.Server(hbmBudget=budget, hostBudget=host, slots=2)
    ids = [srv.addScene(s["loop"]) for s in served]
    st0 = srv.stats()
    assert st0["evictions"] == 1 and st0["resident_scenes"] == 2 and st0["host_bytes"] == one
    srv.wait(srv.submit(ids[1], DT))
    srv.wait(srv.submit(ids[2], DT))
    t0 = srv.submit(ids[0], DT)  # needs an eviction the host budget refuses
    with pytest.raises(pf.PFError, match="host memory"):
        srv.wait(t0)
    with pytest.raises(pf.PFError, match="host memory"):
        srv.evictScene(ids[1])
    srv.wait(srv.submit(ids[1], DT))  # still serving
    st = srv.stats()
    print(f"[serve] stats {st}", flush=True)
    assert st["failed"] == 1 and st["completed"] == 3 and st["evict_failures"] >= 2
    assert st["peak_resident_bytes"] <= budget
    for i in ids:
        srv.removeScene(i)
    steps = [0, 2, 1]
    for s, k in zip(alone, steps):
        for _ in range(k):
            s["loop"].step(DT)
    for k, (a, b) in enumerate(zip(alone, served)):
        assert_same(state(a), state(b), k)
    srv.dispose()
