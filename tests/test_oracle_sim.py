"""The oracle's substep pinned by closed forms, invariants and the reference tests' criteria.

CPU only. FP32 mirror and FP64 restatement of the same schedule.
"""
import math

import numpy as np
import pytest

import oracle as orc
from helpers import oracle_dam_break, rel_err
from particle_fluids import scenes


def poly6(r2, h):
    return max(h * h - r2, 0.0) ** 3 * 315.0 / (64.0 * math.pi * h ** 9)


def test_free_fall_is_semi_implicit_euler():
    # integrate-free-fall.gpu.test.ts:21-44: a lone particle, no colliders, 60 steps of 1/60.
    for prec in ("f32", "f64"):
        s = orc.Scene(1, 0.05, prec)
        s.set_particles(np.array([[0.0, 0.0, 0.0]], np.float32))
        s.set_loop(1, 0, (0.0, -9.81, 0.0), 0.2, 1024)
        ref_v = ref_y = 0.0
        for _ in range(60):
            s.step(1.0 / 60.0)
            ref_v -= 9.81 / 60.0
            ref_y += ref_v / 60.0
        y = s.get()["x"][0, 1]
        assert y < 0
        assert abs(y - ref_y) < 1e-4


def test_isolated_particle_density_and_zero_lambda():
    # lambda.ts:41-66 with only the particle itself in its list: rho = m W(0), C < 0 -> lambda 0, no move.
    r = 0.01
    spacing, h = 2 * r, 4 * r
    for prec in ("f32", "f64"):
        s = orc.Scene(1, r, prec)
        s.set_particles(np.array([[0.1, 0.2, 0.3]], np.float32))
        s.set_loop(1, 1, (0.0, 0.0, 0.0), h, 1024)
        s.add_fluid(0, 1, 1000.0, spacing, h, 1e-4)
        s.step(1.0 / 60.0)
        st = s.get()
        m = 1000.0 * spacing ** 3
        assert st["rho"][0] == pytest.approx(m * poly6(0.0, h), rel=1e-6)
        assert np.allclose(st["x"][0], [0.1, 0.2, 0.3], atol=1e-7)


def test_polar_decomposition_recovers_rotation():
    # softbody-polar-decomp.gpu.test.ts:119-201: A = R S with S symmetric positive definite.
    rng = np.random.default_rng(7)
    for _ in range(50):
        q = rng.standard_normal(4)
        q /= np.linalg.norm(q)
        x, y, z, w = q
        R = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                      [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                      [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])
        B = rng.standard_normal((3, 3))
        S = B @ B.T + 0.5 * np.eye(3)
        A = R @ S
        assert np.abs(orc.polar(A, "f64") - R).max() < 1e-9
        assert np.abs(orc.polar(A, "f32") - R).max() < 1e-4


def lattice(n, spacing, origin):
    g = np.stack(np.meshgrid(*[np.arange(n)] * 3, indexing="ij"), -1).reshape(-1, 3)
    return (origin + g * spacing).astype(np.float32)


def body_scene(pos, rest, compliance, prec, moved=None):
    n = len(pos)
    s = orc.Scene(n, 0.01, prec)
    s.set_particles(pos if moved is None else moved)
    s.set_loop(1, 1, (0.0, 0.0, 0.0), 0.04, 1024)
    c = rest.astype(np.float64).mean(0)
    s.add_body(0, n, compliance, (rest - c).astype(np.float32), np.ones(n, np.float32))
    return s


def test_shape_matching_rigid_motion_is_a_fixed_point():
    # softbody-pass3-apply.gpu.test.ts:187-321: a rigidly moved body needs no correction.
    rest = lattice(5, 0.02, np.array([0.0, 0.0, 0.0]))
    th = 0.3
    Rz = np.array([[math.cos(th), -math.sin(th), 0], [math.sin(th), math.cos(th), 0], [0, 0, 1]])
    c = rest.mean(0)
    moved = ((rest - c) @ Rz.T + c + np.array([0.1, 0.2, -0.05])).astype(np.float32)
    s = body_scene(rest, rest, 1e-6, "f64", moved)
    s.step(1.0 / 60.0)
    # inputs are FP32-rounded, so the moved body is rigid to ~1e-8 m only
    assert np.abs(s.get()["x"] - moved).max() < 1e-7
    cc, RR = s.body_frame(0)
    assert np.abs(RR - Rz).max() < 1e-6


def test_shape_matching_zero_compliance_snaps_to_goal():
    # softbody-pass3-apply.gpu.test.ts:322-391: alpha = 0 moves each particle onto R r + c.
    rest = lattice(4, 0.02, np.array([0.0, 0.0, 0.0]))
    rng = np.random.default_rng(3)
    moved = (rest + rng.normal(0, 0.002, rest.shape)).astype(np.float32)
    s = body_scene(rest, rest, 0.0, "f64", moved)
    s.step(1.0 / 60.0)
    c, R = s.body_frame(0)
    goal = (rest - rest.astype(np.float64).mean(0)) @ R.T + c
    assert np.abs(s.get()["x"] - goal).max() < 1e-12


def test_center_of_mass_of_a_body():
    # softbody-center-of-mass.gpu.test.ts:120-244: c = sum m x / sum m.
    rest = lattice(6, 0.015, np.array([0.3, 0.1, -0.2]))
    s = body_scene(rest, rest, 1e-6, "f32")
    s.step(1.0 / 60.0)
    c, _ = s.body_frame(0)
    assert np.abs(c - rest.astype(np.float64).mean(0)).max() < 1e-5


def test_plane_keeps_a_resting_particle_above_the_floor():
    # collider plane-rest: a particle on the floor under gravity stays at y >= r.
    r = 0.01
    s = orc.Scene(1, r, "f64")
    s.set_particles(np.array([[0.0, r, 0.0]], np.float32))
    s.set_loop(4, 2, (0.0, -9.81, 0.0), 0.04, 1024)
    s.add_plane((0, 1, 0), (0, 0, 0))
    for _ in range(120):
        s.step(1.0 / 60.0)
    st = s.get()
    assert st["x"][0, 1] > r - 1e-4
    assert abs(st["v"][0, 1]) < 1e-2


def test_surface_tension_conserves_momentum():
    # cohesion-newton3.gpu.test.ts:31: pairwise forces sum to zero.
    r = 0.01
    pos = scenes.jitter(lattice(6, 2 * r, np.array([0.0, 0.0, 0.0])), 2 * r, 11, 0.05)
    n = len(pos)
    s = orc.Scene(n, r, "f64")
    s.set_particles(pos)
    s.set_loop(1, 0, (0.0, 0.0, 0.0), 4 * r, 1024)
    s.add_fluid(0, n, 1000.0, 2 * r, 4 * r, 1e-4, surface_tension=0.5)
    s.step(1.0 / 60.0)
    v = s.get()["v"]
    assert np.abs(v).max() > 1e-6
    assert np.abs(v.sum(0)).max() < 1e-9 * np.abs(v).sum()


def test_fp32_mirror_tracks_fp64_for_one_substep():
    pos, r = scenes.dam_break_layout(5000)
    pos = scenes.jitter(pos, 2 * r)
    a = oracle_dam_break(pos, r, 1, 4, "f32")
    b = oracle_dam_break(pos, r, 1, 4, "f64")
    a.step(1.0 / 120.0)
    b.step(1.0 / 120.0)
    sa, sb = a.get(), b.get()
    assert rel_err(sa["x"], sb["x"]) < 1e-6
    assert rel_err(sa["v"], sb["v"]) < 1e-3


def test_dam_break_stays_in_the_container_and_incompressible():
    # mass-conservation.gpu.test.ts:128-129 (no escapes), incompressibility.gpu.test.ts:171-174.
    pos, r = scenes.dam_break_layout(5000)
    pos = scenes.jitter(pos, 2 * r)
    s = oracle_dam_break(pos, r, 2, 4, "f32")
    for _ in range(30):
        s.step(1.0 / 60.0)
    st = s.get()
    x = st["x"]
    lo, hi = np.array(scenes.CONTAINER[:3]), np.array(scenes.CONTAINER[3:])
    assert (x >= lo - 2 * r).all() and (x[:, [0, 2]] <= hi[[0, 2]] + 2 * r).all()
    rho = st["rho"]
    interior = rho > 0.8 * 1000.0
    assert abs(rho[interior].mean() / 1000.0 - 1.0) < 0.05
