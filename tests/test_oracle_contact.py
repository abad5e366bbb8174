"""Particle contacts in the oracle (src/core/contact, loop.ts:346-418).

Pinned by what the passes must do, not by their formulas:
- pair generation against a brute-force search over all pairs (the
  grid walk is not consulted), and the reference's emitter-subset test;
- stabilisation of a resting overlap against the closed form (the pair ends
  exactly 2r apart, split by inverse mass, with no velocity created);
- the reference's head-on test (no interpenetration, momentum kept), its
  overflow test, and its pile / incline tests;
- static friction cancelling the slip of a pressed particle and kinetic
  friction slowing it by mu_k g dt (Coulomb);
- a scene whose only particles are fluid runs unchanged with contacts on
  (fluids never emit).
"""
import numpy as np
import pytest

from helpers import brute_force_pairs, oracle_contact_scene, oracle_dam_break
from particle_fluids import scenes


def unordered(pairs):
    return {(int(min(a, b)), int(max(a, b))) for a, b in pairs}


def lattice(side, spacing, y0=0.5):
    g = np.stack(np.meshgrid(*(np.arange(side),) * 3, indexing="ij"), -1).reshape(-1, 3).astype(np.float64)
    g = (g - (side - 1) / 2) * spacing
    g[:, 1] += y0 + (side - 1) / 2 * spacing
    return g


def test_generation_matches_brute_force():
    rng = np.random.default_rng(7)
    r = 0.02
    n = 400
    pos = rng.uniform(-0.15, 0.15, (n, 3))
    w = np.where(rng.random(n) < 0.15, 0.0, 1.0).astype(np.float32)
    groups = rng.integers(0, 4, n).astype(np.uint32)
    quiet = [(0, 60), (200, 40)]
    emitter = np.ones(n, bool)
    for a, c in quiet:
        emitter[a:a + c] = False
    s = oracle_contact_scene(pos, r, 1, 0, gravity=(0, 0, 0), w=w, groups=groups, quiet=quiet)
    s.step(1 / 60)
    count, pairs = s.contacts()
    # No gravity, no velocity: the predictions the search ran on are the
    # uploaded positions.
    expect = brute_force_pairs(pos, w, groups, r, emitter)
    assert count == len(pairs) == len(expect) > 100
    assert unordered(pairs) == expect
    for i, j in pairs:
        # stored from the emitter's side, the lower index when both emit
        assert emitter[i]
        assert not emitter[j] or i < j


def test_emitter_subset_finds_the_same_pairs():
    """contact-generate-emitters.gpu.test.ts: a 200-particle column on a pinned
    60-particle slab, searched by every particle and by the slab only."""
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

    def run(quiet):
        s = oracle_contact_scene(pos, r, 1, 1, w=w, groups=grp, max_contacts=8 * len(pos), quiet=quiet)
        s.step(1 / 60)
        return s.contacts()[1]

    allp = run(())
    emit = run([(0, col)])
    assert len(emit) > 0
    assert unordered(allp) == unordered(emit)
    assert all(i >= col for i, _ in emit)
    assert all(i < j for i, j in allp)


@pytest.mark.parametrize("w2", [1.0, 3.0, 0.0])
def test_stabilize_closed_form(w2):
    """An overlap present at the start of a substep is removed before the
    solve, split by inverse mass, moving x and x* alike (solve.ts:134-181):
    the pair ends 2r apart and no velocity appears."""
    r = 0.05
    d0 = 0.07
    pos = np.array([[0.0, 0.0, 0.0], [d0, 0.0, 0.0]])
    w = np.array([1.0, w2], np.float32)
    s = oracle_contact_scene(pos, r, 1, 0, gravity=(0, 0, 0), w=w, precision="f64")
    s.step(1 / 240)
    st = s.get()
    depth = 2 * r - d0
    ws = 1.0 + w2
    assert abs(st["x"][0, 0] - (-depth * 1.0 / ws)) < 1e-8
    assert abs(st["x"][1, 0] - (d0 + depth * w2 / ws)) < 1e-8
    assert abs((st["x"][1, 0] - st["x"][0, 0]) - 2 * r) < 1e-8
    assert np.all(st["v"] == 0.0)


@pytest.mark.parametrize("precision", ["f32", "f64"])
def test_headon_no_penetration_and_momentum(precision):
    """contact-headon.gpu.test.ts: two particles of r = 0.5 meet at +-0.5 m/s
    with no friction and no gravity."""
    r = 0.5
    pos = np.array([[-0.75, 0, 0], [0.75, 0, 0]], np.float64)
    vel = np.array([[0.5, 0, 0], [-0.5, 0, 0]], np.float64)
    s = oracle_contact_scene(pos, r, 4, 8, gravity=(0, 0, 0), vel=vel, mus=0.0, muk=0.0, max_contacts=16,
                             precision=precision)
    dmin = np.inf
    for _ in range(120):
        s.step(1 / 60)
        x = s.get()["x"]
        dmin = min(dmin, np.linalg.norm(x[0] - x[1]))
    v = s.get()["v"]
    assert dmin > 2 * r - 1e-3
    assert abs(v[0, 0] + v[1, 0]) < 1e-4


def test_overflow_counts_past_the_cap():
    """contact-overflow.gpu.test.ts: 125 packed particles, maxContacts 16."""
    r = 0.05
    pos = lattice(5, 2 * r * 1.01)
    s = oracle_contact_scene(pos, r, 2, 2, mus=0.4, muk=0.3, max_contacts=16)
    s.step(1 / 60)
    count, pairs = s.contacts()
    assert count > 16 and len(pairs) == 16
    for _ in range(10):
        s.step(1 / 60)
    assert np.all(np.isfinite(s.get()["x"]))


def test_pile_on_floor_never_overlaps_and_settles():
    """contact-non-interpenetration and contact-small-cluster: a 3^3 pile
    dropped on a plane keeps every pair at least 2r - 2e-3 apart and settles
    with its bottom layer at y ~ r."""
    r = 0.05
    pos = lattice(3, 2 * r * 1.05)
    s = oracle_contact_scene(pos, r, 4, 4, mus=0.4, muk=0.3, max_contacts=256)
    s.add_plane((0, 1, 0), (0, 0, 0))
    worst = 0.0
    for _ in range(90):
        s.step(1 / 60)
        x = s.get()["x"]
        d = np.linalg.norm(x[:, None] - x[None], axis=-1) + np.eye(len(x)) * 10
        worst = max(worst, 2 * r - d.min())
    st = s.get()
    assert worst < 2e-3
    assert st["x"][:, 1].min() > r - 2e-3
    assert np.linalg.norm(st["v"], axis=1).max() < 1.0


def pressed_pair(vx, mus, muk, precision="f64"):
    """Particle 1 resting exactly on pinned particle 0 (distance 2r), moving
    sideways at vx under gravity."""
    r = 0.05
    pos = np.array([[0.0, 0.0, 0.0], [0.0, 2 * r, 0.0]])
    vel = np.array([[0.0, 0.0, 0.0], [vx, 0.0, 0.0]])
    w = np.array([0.0, 1.0], np.float32)
    return oracle_contact_scene(pos, r, 1, 1, vel=vel, w=w, mus=mus, muk=muk, precision=precision)


def test_static_friction_cancels_the_slip():
    """One solve of a pressed particle, in closed form: the normal push takes
    it back to distance 2r along the contact normal, and inside the friction
    cone the tangential part of its slip is removed whole (solve.ts:70-92),
    so it barely moves sideways; without static friction it keeps vx dt.
    Tolerance: two ticks of the positions accumulator (10 / 2^30 m)."""
    r, g, dt, vx = 0.05, 9.81, 1 / 240, 0.01
    xs = np.array([vx * dt, 2 * r - g * dt * dt, 0.0])  # prediction after gravity
    d = np.linalg.norm(xs)
    n = xs / d
    push = n * (2 * r - d)
    slip = xs - np.array([0.0, 2 * r, 0.0])
    tang = slip - n * slip.dot(n)
    tick = 10.0 / 2 ** 30
    for mus, want in ((10.0, xs + push - tang), (0.0, xs + push)):
        s = pressed_pair(vx, mus=mus, muk=0.0)
        s.step(dt)
        got = s.get()["x"][1]
        assert np.abs(got - want).max() < 2 * tick, (mus, got, want)
    assert abs((xs + push - tang)[0]) < 1e-2 * vx * dt


def test_kinetic_friction_is_coulomb():
    """A pressed particle sliding sideways, one substep in closed form: the
    normal solve sets lambda_n = depth (w = 1, lambda in 2^-20 ticks), and
    kinetic friction then removes min(mu_k lambda_n / dt, |v_t|) of its
    tangential velocity (solve.ts:147-182): about mu_k g dt, Coulomb.
    Tolerance: the positions accumulator tick over dt."""
    r, g, dt, vx, muk = 0.05, 9.81, 1 / 240, 0.05, 0.5
    x0 = np.array([0.0, 2 * r, 0.0])
    xs = np.array([vx * dt, 2 * r - g * dt * dt, 0.0])
    d = np.linalg.norm(xs)
    n = xs / d
    depth = 2 * r - d
    v1 = (xs + n * depth - x0) / dt
    vt = v1 - n * n.dot(v1)
    lam = np.floor(depth * 2 ** 20) / 2 ** 20
    change = min(muk * lam / dt, np.linalg.norm(vt))
    want = v1 - vt / np.linalg.norm(vt) * change
    s = pressed_pair(vx, mus=0.0, muk=muk)
    s.step(dt)
    got = s.get()["v"][1]
    assert np.abs(got - want).max() < 2 * (10.0 / 2 ** 30) / dt, (got, want)
    assert abs(change - muk * g * dt) < 0.02 * muk * g * dt


@pytest.mark.parametrize("kind", ["static", "kinetic"])
def test_two_balls_on_incline(kind):
    """contact-two-ball-inclined.gpu.test.ts: two balls on a plane under
    gravity tilted atan(mu_s) -+ 3 degrees."""
    r, mus, muk, g = 0.05, 0.6, 0.5, 9.81
    theta = np.arctan(mus) + (-3 if kind == "static" else 3) * np.pi / 180
    seconds = 5.0 if kind == "static" else 2.0
    pos = np.array([[-0.2, r, 0], [0.2, r, 0]])
    s = oracle_contact_scene(pos, r, 4, 4, gravity=(g * np.sin(theta), -g * np.cos(theta), 0), mus=mus, muk=muk,
                             max_contacts=8)
    s.add_plane((0, 1, 0), (0, 0, 0), mus, muk)
    for _ in range(int(np.ceil(seconds * 60))):
        s.step(1 / 60)
    dx = s.get()["x"][:, 0] - pos[:, 0]
    if kind == "static":
        assert np.abs(dx).max() < 1e-3
    else:
        a = g * (np.sin(theta) - muk * np.cos(theta))
        want = 0.5 * a * seconds * seconds
        assert np.all(dx > 0.85 * want) and np.all(dx < 1.15 * want), (dx, want)


def test_fluid_only_scene_unchanged_by_contacts():
    """Fluids keep their own particles apart (FluidSystem.ts:308): with every
    particle in a fluid nothing emits, and the step matches the run without
    contacts."""
    pos, r = scenes.dam_break_layout(1500)
    pos = scenes.jitter(pos, 2 * r)
    a = oracle_dam_break(pos, r, 2, 2)
    b = oracle_dam_break(pos, r, 2, 2)
    b.set_contact(0.3, 0.2)
    for _ in range(3):
        a.step(1 / 60)
        b.step(1 / 60)
    assert b.contacts()[0] == 0
    sa, sb = a.get(), b.get()
    assert np.array_equal(sa["x"], sb["x"]) and np.array_equal(sa["v"], sb["v"])
