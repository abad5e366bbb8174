"""Particle contacts on the GPU (src/core/contact, loop.ts:346-418) against
the FP32 oracle and the reference's contact scenarios.

Scenes of contacts and colliders only are bit-identical with the oracle: the
fixed-point sums make the scatter order irrelevant. Scenes with soft bodies
carry the shape-matching reduction order, so they are held to the 1e-5
relative tolerance of the other soft-body scenes."""
import numpy as np
import pytest

import particle_fluids as pf
from particle_fluids import scenes
from helpers import brute_force_pairs, floating_cubes_layout, oracle_contact_scene, oracle_mixed_scene, rel_err

pytestmark = pytest.mark.gpu


def lattice(side, spacing, y0=0.5):
    g = np.stack(np.meshgrid(*(np.arange(side),) * 3, indexing="ij"), -1).reshape(-1, 3).astype(np.float64)
    g = (g - (side - 1) / 2) * spacing
    g[:, 1] += y0 + (side - 1) / 2 * spacing
    return g


def gpu_contact_scene(pos, r, substeps, iterations, gravity=(0.0, -9.81, 0.0), vel=None, w=None, groups=None,
                      contact=True, quiet=(), plane=None):
    ps = pf.ParticleSystem(len(pos), r)
    ps.uploadParticles(pos, vel, w, groups)
    cols = []
    if plane is not None:
        walls = pf.PrimitiveSet(ps)
        walls.addPlane(plane[0], plane[1], muS=plane[2], muK=plane[3])
        cols.append(walls)
    mats = [pf.NoSelfContacts(q) for q in quiet]
    loop = pf.SimLoop(ps, substeps=substeps, iterations=iterations, gravity=gravity, materials=mats,
                      colliders=cols, contact=contact)
    return ps, loop, cols


def unordered(pairs):
    p = np.asarray(pairs).reshape(-1, 2)
    return {(int(min(a, b)), int(max(a, b))) for a, b in p}


def xv(ps):
    s = ps.readback()
    return s.positions[:, :3].astype(np.float64), s.velocities[:, :3].astype(np.float64)


def test_pair_search_matches_brute_force_and_oracle():
    rng = np.random.default_rng(7)
    r, n = 0.02, 400
    pos = rng.uniform(-0.15, 0.15, (n, 3))
    w = np.where(rng.random(n) < 0.15, 0.0, 1.0).astype(np.float32)
    groups = rng.integers(0, 4, n).astype(np.uint32)
    quiet = [(0, 60), (200, 40)]
    emitter = np.ones(n, bool)
    for a, c in quiet:
        emitter[a:a + c] = False
    print("[contact] pair search, 400 particles, 2 quiet ranges", flush=True)
    ps, loop, _ = gpu_contact_scene(pos, r, 1, 0, gravity=(0, 0, 0), w=w, groups=groups, quiet=quiet)
    loop.step(1 / 60)
    count = loop.contacts.readbackCount()
    stored, pairs = loop.contacts.readbackPairs()
    expect = brute_force_pairs(pos, w, groups, r, emitter)
    assert count == stored == len(expect) > 100
    assert unordered(pairs) == expect
    for i, j in pairs.reshape(-1, 2):
        assert emitter[i] and (not emitter[j] or i < j)
    o = oracle_contact_scene(pos, r, 1, 0, gravity=(0, 0, 0), w=w, groups=groups, quiet=quiet)
    o.step(1 / 60)
    oc, op = o.contacts()
    assert oc == count and unordered(op) == unordered(pairs)
    loop.dispose()
    ps.dispose()


def test_emitter_subset_finds_the_same_pairs():
    """contact-generate-emitters.gpu.test.ts through a material that only
    declares noSelfContacts."""
    r = 0.025
    sp = 2 * r
    col, slab = 200, 60
    side = int(np.ceil(np.cbrt(col)))
    pos, w, grp = [], [], []
    for i in range(col):
        x, y, z = i % side, (i // side) % side, i // (side * side)
        pos.append([-side * sp * 0.5 + (x + 0.5) * sp, (y + 0.5) * sp, -side * sp * 0.5 + (z + 0.5) * sp])
        w.append(1.0)
        grp.append(1)
    ss = int(np.ceil(np.sqrt(slab)))
    for i in range(slab):
        x, z = i % ss, i // ss
        pos.append([-ss * sp * 0.5 + (x + 0.5) * sp, -sp * 0.5, -ss * sp * 0.5 + (z + 0.5) * sp])
        w.append(0.0)
        grp.append(2)
    pos, w, grp = np.array(pos), np.array(w, np.float32), np.array(grp, np.uint32)
    print("[contact] emitter subset", flush=True)

    def run(quiet):
        ps, loop, _ = gpu_contact_scene(pos, r, 1, 1, w=w, groups=grp, contact={"maxContacts": 8 * len(pos)},
                                        quiet=quiet)
        loop.step(1 / 60)
        pairs = loop.contacts.readbackPairs()[1].reshape(-1, 2)
        loop.dispose()
        ps.dispose()
        return pairs

    allp = run(())
    emit = run([(0, col)])
    assert len(emit) > 0
    assert unordered(allp) == unordered(emit)
    assert np.all(emit[:, 0] >= col)
    assert np.all(allp[:, 0] < allp[:, 1])


def test_pile_on_floor_bit_identical_to_oracle():
    """contact-non-interpenetration's 3^3 pile on a plane: every step equals
    the FP32 oracle bit for bit, and the pile keeps its pairs apart."""
    r = 0.05
    pos = lattice(3, 2 * r * 1.05)
    print("[contact] pile on floor vs oracle, 27 particles, 30 steps", flush=True)
    ps, loop, cols = gpu_contact_scene(pos, r, 4, 4, contact={"maxContacts": 256, "muS": 0.4, "muK": 0.3},
                                       plane=((0, 1, 0), (0, 0, 0), 0.5, 0.4))
    o = oracle_contact_scene(pos, r, 4, 4, mus=0.4, muk=0.3, max_contacts=256)
    o.add_plane((0, 1, 0), (0, 0, 0))
    worst = 0.0
    for k in range(30):
        loop.step(1 / 60)
        o.step(1 / 60)
        x, v = xv(ps)
        ref = o.get()
        assert np.array_equal(x, ref["x"]) and np.array_equal(v, ref["v"]), k
        d = np.linalg.norm(x[:, None] - x[None], axis=-1) + np.eye(len(x)) * 10
        worst = max(worst, 2 * r - d.min())
    assert worst < 2e-3
    assert loop.contacts.readbackCount() == o.contacts()[0]
    loop.dispose()
    cols[0].dispose()
    ps.dispose()


def test_headon_bit_identical_no_penetration():
    """contact-headon.gpu.test.ts on the GPU, checked against the oracle
    every frame."""
    r = 0.5
    pos = np.array([[-0.75, 0, 0], [0.75, 0, 0]])
    vel = np.array([[0.5, 0, 0], [-0.5, 0, 0]])
    print("[contact] head-on, 120 frames", flush=True)
    ps, loop, _ = gpu_contact_scene(pos, r, 4, 8, gravity=(0, 0, 0), vel=vel,
                                    contact={"maxContacts": 16, "muS": 0, "muK": 0})
    o = oracle_contact_scene(pos, r, 4, 8, gravity=(0, 0, 0), vel=vel, mus=0.0, muk=0.0, max_contacts=16)
    dmin = np.inf
    for k in range(120):
        loop.step(1 / 60)
        o.step(1 / 60)
        x, v = xv(ps)
        ref = o.get()
        assert np.array_equal(x, ref["x"]) and np.array_equal(v, ref["v"]), k
        dmin = min(dmin, np.linalg.norm(x[0] - x[1]))
    assert dmin > 2 * r - 1e-3
    assert abs(v[0, 0] + v[1, 0]) < 1e-4
    loop.dispose()
    ps.dispose()


def test_overflow_is_reported_and_stays_finite():
    """contact-overflow.gpu.test.ts: 125 packed particles, maxContacts 16."""
    r = 0.05
    pos = lattice(5, 2 * r * 1.01)
    print("[contact] overflow, cap 16", flush=True)
    ps, loop, _ = gpu_contact_scene(pos, r, 2, 2, contact={"maxContacts": 16, "muS": 0.4, "muK": 0.3})
    loop.step(1 / 60)
    assert loop.contacts.readbackCount() > 16
    assert loop.contacts.readbackPairs()[0] == 16
    assert loop.readbackOverflow().contacts
    for _ in range(10):
        loop.step(1 / 60)
    assert np.all(np.isfinite(ps.readback().positions))
    loop.dispose()
    ps.dispose()


@pytest.mark.parametrize("kind", ["static", "kinetic"])
def test_two_balls_on_incline(kind):
    """contact-two-ball-inclined.gpu.test.ts: bit-identical with the oracle,
    and the reference's bounds."""
    r, mus, muk, g = 0.05, 0.6, 0.5, 9.81
    theta = np.arctan(mus) + (-3 if kind == "static" else 3) * np.pi / 180
    seconds = 5.0 if kind == "static" else 2.0
    pos = np.array([[-0.2, r, 0], [0.2, r, 0]])
    grav = (g * np.sin(theta), -g * np.cos(theta), 0)
    print(f"[contact] two balls on incline, {kind}", flush=True)
    ps, loop, cols = gpu_contact_scene(pos, r, 4, 4, gravity=grav, contact={"maxContacts": 8, "muS": mus, "muK": muk},
                                       plane=((0, 1, 0), (0, 0, 0), mus, muk))
    o = oracle_contact_scene(pos, r, 4, 4, gravity=grav, mus=mus, muk=muk, max_contacts=8)
    o.add_plane((0, 1, 0), (0, 0, 0), mus, muk)
    for _ in range(int(np.ceil(seconds * 60))):
        loop.step(1 / 60)
        o.step(1 / 60)
    x, v = xv(ps)
    ref = o.get()
    assert np.array_equal(x, ref["x"]) and np.array_equal(v, ref["v"])
    dx = x[:, 0] - pos[:, 0]
    if kind == "static":
        assert np.abs(dx).max() < 1e-3
    else:
        a = g * (np.sin(theta) - muk * np.cos(theta))
        want = 0.5 * a * seconds * seconds
        assert np.all(dx > 0.85 * want) and np.all(dx < 1.15 * want)
    loop.dispose()
    cols[0].dispose()
    ps.dispose()


def test_cubes_on_water_with_contacts_match_oracle():
    """Two soft cubes resting on the dam-break water, overlapping it and each
    other, with Simulation's contacts (muS 0.3, muK 0.2). The stabilisation
    moves water particles between the sort and the neighbour build, which
    exercises the grid's sorted copy of the predictions."""
    print("[contact] cubes on water vs oracle: budget 3000, 2 cubes of 125", flush=True)
    sc = scenes.build_layout_scene(floating_cubes_layout())
    assert sc["loop"].contacts is not None
    H = sc["loop"].info()["hashTableSize"]
    o = oracle_mixed_scene(sc["layout"], 2, 4, H)
    errs, counts = [], []
    for _ in range(4):
        sc["loop"].step(1.0 / 60.0)
        o.step(1.0 / 60.0)
        s = sc["ps"].readback()
        ref = o.get()
        errs.append((rel_err(s.positions[:, :3], ref["x"]), rel_err(s.velocities[:, :3], ref["v"])))
        counts.append((sc["loop"].contacts.readbackCount(), o.contacts()[0]))
    print(f"[contact] rel err per step (x, v) {errs}; contacts (gpu, oracle) {counts}", flush=True)
    assert counts[0][0] == counts[0][1] > 0
    assert all(abs(a - b) <= 2 for a, b in counts)
    for ex, ev in errs:
        assert ex <= 1e-5 and ev <= 1e-4, errs
    scenes.dispose_scene(sc)


@pytest.mark.parametrize("graph", [False, True])
def test_batch_with_contact_scenes_matches_each_alone(graph):
    """Scenes with contacts (cubes on water, cubes in the tank extension)
    beside one without, in one launch: each ends bit-identical to its run
    alone."""
    spec = [(3000, 0), (4000, 2), (2500, 1)]
    print(f"[contact] batch vs alone, graph={graph}, {spec} + cubes on water", flush=True)

    def build():
        out = [scenes.build_mixed_scene(b, nb, seed=200 + i) for i, (b, nb) in enumerate(spec)]
        out.insert(1, scenes.build_layout_scene(floating_cubes_layout(2000, seed=5)))
        return out

    alone, mixed = build(), build()
    batch = pf.Batch([s["loop"] for s in mixed])
    batch.useGraph(graph)
    for _ in range(3):
        for s in alone:
            s["loop"].step(1.0 / 60.0)
        batch.step(1.0 / 60.0)
    for i, (a, b) in enumerate(zip(alone, mixed)):
        sa, sb = a["ps"].readback(), b["ps"].readback()
        assert np.array_equal(sa.positions, sb.positions), i
        assert np.array_equal(sa.velocities, sb.velocities), i
        if a["loop"].contacts is not None:
            assert a["loop"].contacts.readbackCount() == b["loop"].contacts.readbackCount()
    batch.dispose()
    for s in alone + mixed:
        scenes.dispose_scene(s)
