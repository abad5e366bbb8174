"""The oracle's fluid boundaries (boundaryVolume.ts:26-52, positionDelta.ts:54-69)
pinned by closed forms: an isolated boundary particle, a lattice interior
summed by brute force, and momentum conservation of the boundary reaction."""
import math

import numpy as np

import oracle as orc
from helpers import boundary_layout, oracle_boundary_scene


def poly6(r2, h):
    return max(h * h - r2, 0.0) ** 3 * 315.0 / (64.0 * math.pi * h ** 9)


def test_isolated_boundary_particle_volume():
    # one fluid particle far away, one static boundary particle: V = 1 / W(0)
    r = 0.01
    h = 4 * r
    pos = np.array([[0.0, 0.5, 0.0], [0.2, 0.1, 0.1]], np.float32)
    for prec, tol in (("f64", 1e-12), ("f32", 1e-6)):
        s = orc.Scene(2, r, prec)
        s.set_particles(pos, w=np.array([1.0, 0.0], np.float32))
        s.set_loop(1, 1, (0.0, 0.0, 0.0), h, 64)
        s.add_fluid(0, 1, 1000.0, 2 * r, h, 1e-4)
        s.add_boundary(0, 1, 1, dynamic=False)
        s.step(1e-3)
        v = s.boundary_volumes()[1]
        assert abs(v - 64.0 * math.pi * h ** 3 / 315.0) <= tol * v, (prec, v)


def test_lattice_interior_volume_matches_brute_force():
    # 9^3 static boundary lattice at spacing h/2: the centre particle's sum
    # runs over every lattice offset inside h
    r = 0.01
    h = 4 * r
    sp = h / 2
    g = np.arange(-4, 5) * sp
    lat = np.array([(x, y, z) for x in g for y in g for z in g], np.float64) + np.array([0.0, 0.4, 0.0])
    pos = np.concatenate([[[0.3, 0.7, 0.2]], lat]).astype(np.float32)
    n = len(pos)
    centre = 1 + (9 * 9 * 9) // 2
    d = pos[1:].astype(np.float64) - pos[centre].astype(np.float64)
    ref = 1.0 / sum(poly6(float((e * e).sum()), h) for e in d)
    for prec, tol in (("f64", 1e-9), ("f32", 1e-5)):
        s = orc.Scene(n, r, prec)
        w = np.zeros(n, np.float32)
        w[0] = 1.0
        s.set_particles(pos, w=w)
        s.set_loop(1, 1, (0.0, 0.0, 0.0), h, 4096)
        s.add_fluid(0, 1, 1000.0, 2 * r, h, 1e-4)
        s.add_boundary(0, 1, n - 1, dynamic=False)
        s.step(1e-3)
        v = s.boundary_volumes()[centre]
        assert abs(v - ref) <= tol * ref, (prec, v, ref)


def test_static_and_dynamic_volumes_agree_for_a_pinned_plate():
    pos, w, nf, nb = boundary_layout(plate_w=0.0)
    r = 0.01
    a = oracle_boundary_scene(pos, w, nf, nb, r, dynamic=False, precision="f64")
    b = oracle_boundary_scene(pos, w, nf, nb, r, dynamic=True, precision="f64")
    a.step(1 / 60)
    b.step(1 / 60)
    va, vb = a.boundary_volumes()[nf:], b.boundary_volumes()[nf:]
    assert np.all(va > 0)
    assert np.allclose(va, vb, rtol=1e-12)


def test_reaction_conserves_momentum():
    # no gravity, walls or effects: the fluid sits compressed against a free
    # plate, so every push it gets has an equal and opposite reaction
    r = 0.01
    pos, w, nf, nb = boundary_layout(radius=r)
    pos[:nf, 1] -= 0.035  # press the block into the plate
    m_fluid = 1000.0 * (2 * r) ** 3
    w[nf:] = 1.0 / (0.5 * m_fluid)
    s = oracle_boundary_scene(pos, w, nf, nb, r, dynamic=True, substeps=1, iterations=4, precision="f64",
                              gravity=(0.0, 0.0, 0.0), effects={}, planes=False)
    s.step(1 / 120)
    v = s.get()["v"]
    m = np.concatenate([np.full(nf, m_fluid), np.full(nb, 0.5 * m_fluid)])
    p = (m[:, None] * v).sum(0)
    scale = (m[:, None] * np.abs(v)).sum()
    assert scale > 0
    assert np.abs(p).max() <= 1e-6 * scale, (p, scale)
