"""Launch shapes: the persistent form (a capped grid whose workgroups walk the
tile list) gives results bit-identical to one workgroup per tile, for a loop,
a batch and a server, in every kernel family of the substep: the sort, the
neighbour build and constraint passes, surface tension, boundaries, the
viscosity solver, walls, contacts and the grid-free prediction."""
import numpy as np
import pytest

import particle_fluids as pf
from particle_fluids import scenes
from helpers import boundary_layout
from test_gpu_boundary import gpu_boundary_scene
from test_gpu_viscosity import thick_dam_break

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


@pytest.mark.parametrize("shape,graph", [(8, False), (64, True), (4096, True)])
def test_persistent_loop_matches_one_workgroup_per_tile(shape, graph):
    # 20,000 particles are 79 tiles: 8 workgroups walk ten tiles each, 64 walk
    # one or two, and 4096 covers the list (one workgroup per tile again)
    print(f"[launch] dam break, shape {shape}, graph {graph}", flush=True)
    a = scenes.build_dam_break(20000, seed=7)
    b = scenes.build_dam_break(20000, seed=7, use_graph=graph)
    b["loop"].setLaunchShape(shape)
    for _ in range(3):
        a["loop"].step(DT)
        b["loop"].step(DT)
    assert_same(state(a), state(b), shape)


def test_persistent_batch_of_mixed_scenes_matches_alone():
    # soft cubes with particle contacts beside pure fluids: contact search and
    # stabilisation, walls and advect run as tile kernels of the batch too
    spec = [(3000, 0), (6000, 2), (4000, 3), (9000, 0), (2500, 1)]
    print(f"[launch] batch {spec}, shape 24", flush=True)
    alone = [scenes.build_mixed_scene(b, nb, seed=300 + i) for i, (b, nb) in enumerate(spec)]
    mixed = [scenes.build_mixed_scene(b, nb, seed=300 + i) for i, (b, nb) in enumerate(spec)]
    batch = pf.Batch([m["loop"] for m in mixed])
    batch.useGraph(True)
    batch.setLaunchShape(24)
    for _ in range(3):
        for s in alone:
            s["loop"].step(DT)
        batch.step(DT)
    for i, (s, m) in enumerate(zip(alone, mixed)):
        assert_same(state(s), state(m), i)


def test_persistent_boundary_and_viscosity_scenes():
    print("[launch] dynamic boundary and viscosity solver, shape 8", flush=True)
    r = 0.01
    pos, w, nf, nb = boundary_layout(radius=r, plate_w=0.0)
    a = gpu_boundary_scene(pos, w, nf, nb, r, True)
    b = gpu_boundary_scene(pos, w, nf, nb, r, True)
    b["loop"].setLaunchShape(8)
    ta = thick_dam_break(6000, 20.0, 9, seed=11)
    tb = thick_dam_break(6000, 20.0, 9, seed=11)
    tb["loop"].setLaunchShape(8)
    for _ in range(3):
        for s in (a, b, ta, tb):
            s["loop"].step(DT)
    assert_same(state(a), state(b), "boundary")
    assert_same(state(ta), state(tb), "viscosity")


def soft_only_scene(seed):
    """The cubes of a mixed layout without its water, on the tank's planes and
    without contacts: a scene with no neighbour grid."""
    lay = scenes.mixed_layout(3000, 3, seed)
    nf = lay["fluid_count"]
    pos, inv = lay["positions"][nf:], lay["inv_mass"][nf:]
    bodies = [dict(b, range=(b["range"][0] - nf, b["range"][1])) for b in lay["bodies"]]
    ps = pf.ParticleSystem(len(pos), lay["radius"])
    ps.uploadParticles(pos, invMass=inv)
    soft = pf.SoftbodySystem(ps, bodies)
    walls = pf.PrimitiveSet(ps)
    mn, mx = lay["container"][:3], lay["container"][3:]
    for n, which in scenes.PLANES:
        walls.addPlane(n, mn if which == "min" else mx)
    loop = pf.SimLoop(ps, substeps=2, iterations=4, materials=[soft], colliders=[walls])
    return dict(ps=ps, soft=soft, walls=walls, loop=loop)


def test_persistent_grid_free_soft_scene():
    # cubes without fluid or contacts: the prediction pass without a grid
    print("[launch] soft cubes without contacts, shape 8", flush=True)
    a = soft_only_scene(21)
    b = soft_only_scene(21)
    b["loop"].setLaunchShape(8)
    for _ in range(3):
        a["loop"].step(DT)
        b["loop"].step(DT)
    assert_same(state(a), state(b), "soft")


def test_persistent_server_matches_alone():
    spec = [(3000, 0), (6000, 2), (4000, 1)]
    print(f"[launch] server {spec}, shape 16", flush=True)
    alone = [scenes.build_mixed_scene(b, nb, seed=400 + i) for i, (b, nb) in enumerate(spec)]
    served = [scenes.build_mixed_scene(b, nb, seed=400 + i, evictable=True) for i, (b, nb) in enumerate(spec)]
    srv = pf.Server()
    srv.setLaunchShape(16)
    ids = [srv.addScene(s["loop"]) for s in served]
    for _ in range(2):
        for i in ids:
            srv.submit(i, DT)
    srv.drain()
    for i in ids:
        srv.removeScene(i)
    srv.dispose()
    for s in alone:
        for _ in range(2):
            s["loop"].step(DT)
    for i, (s, m) in enumerate(zip(alone, served)):
        assert_same(state(s), state(m), i)


def test_launch_autotune_and_validation():
    sc = scenes.build_dam_break(20000, seed=8, use_graph=True)
    twin = scenes.build_dam_break(20000, seed=8)
    best, times = sc["loop"].autotuneLaunch(DT, candidates=(0, 64), steps=1)
    print(f"[launch] loop autotune best {best} {times}", flush=True)
    assert best in (0, 64) and set(times) == {0, 64} and all(t > 0 for t in times.values())
    # 2 sweeps x 2 candidates x (1 capture + 1 timed) steps
    for _ in range(8):
        twin["loop"].step(DT)
    assert_same(state(twin), state(sc), "autotuned")
    with pytest.raises(pf.PFError, match="multiple of 8"):
        sc["loop"].setLaunchShape(12)
    batch = pf.Batch([sc["loop"]])
    with pytest.raises(pf.PFError, match="multiple of 8"):
        batch.autotuneLaunch(DT, candidates=(0, 100))
    batch.dispose()
