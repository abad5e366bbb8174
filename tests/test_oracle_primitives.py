"""Primitive colliders of the oracle (primitives.ts:85-217, collider.ts:53-118)
against closed forms: one contact solve lands a particle at the particle
radius from the surface, along the shape's normal. Residuals: FP32 storage of
the shape (as uploaded) and the positions accumulator's 1/107374182 m tick."""
import math

import numpy as np
import pytest

import oracle as orc
from helpers import PosedPrimitive, qapply

R_P = 0.01
TOL = 3e-8  # f32 shape data + one fixed-point tick


def one_step(points, add, precision="f64", substeps=1, iterations=1):
    pts = np.asarray(points, np.float64)
    s = orc.Scene(len(pts), R_P, precision)
    s.set_particles(pts, None, np.ones(len(pts)))
    s.set_loop(substeps, iterations, (0.0, 0.0, 0.0), 0.0, 1)
    add(s)
    s.step(1.0 / 60.0)
    return s.get()["x"], s.get()["v"]


def f32(v):
    return float(np.float32(v))


def test_sphere_contact_lands_on_the_surface():
    c, rad = np.array([0.1, 0.2, -0.05]), 0.045
    dirs = np.array([[0, 1, 0], [1, 0, 0], [0.6, 0.8, 0], [0, -0.6, 0.8], [-0.48, 0.6, 0.64]], np.float64)
    pts = c + dirs * 0.048  # inside the reach (radius + particle radius)
    x, v = one_step(pts, lambda s: s.add_sphere(c, rad))
    dist = np.linalg.norm(x - c.astype(np.float32).astype(np.float64), axis=1)
    assert np.all(np.abs(dist - (f32(rad) + R_P)) < TOL), dist - (rad + R_P)
    # pushed radially: the direction from the centre is unchanged
    got = (x - c) / dist[:, None]
    assert np.abs(got - dirs).max() < 1e-6


def test_inverted_sphere_keeps_particles_inside():
    c, rad = np.zeros(3), 0.3
    pts = np.array([[0.295, 0, 0], [0, -0.298, 0], [0.1, 0.1, 0.05]])
    x, _ = one_step(pts, lambda s: s.add_sphere(c, rad, invert=True))
    dist = np.linalg.norm(x, axis=1)
    assert abs(dist[0] - (f32(rad) - R_P)) < TOL and abs(dist[1] - (f32(rad) - R_P)) < TOL
    assert np.allclose(x[2], pts[2])  # 0.15 from the wall: untouched


@pytest.mark.parametrize("angle", [0.0, 30.0])
def test_box_face_contact(angle):
    c, he = np.array([0.0, 0.1, 0.0]), np.array([0.1, 0.05, 0.08])
    q = np.array([0, 0, math.sin(math.radians(angle) / 2), math.cos(math.radians(angle) / 2)])
    nx = qapply(q, [1.0, 0.0, 0.0])
    ny = qapply(q, [0.0, 1.0, 0.0])
    # just inside the +x face, and just outside the +y face within reach
    pts = np.array([c + nx * (he[0] - 0.004) + ny * 0.01, c + ny * (he[1] + 0.005) - nx * 0.02])
    x, _ = one_step(pts, lambda s: s.add_box(c, he, q))
    d0 = np.dot(x[0] - c, nx)
    d1 = np.dot(x[1] - c, ny)
    assert abs(d0 - (he[0] + R_P)) < 1e-6 and abs(d1 - (he[1] + R_P)) < 1e-6, (d0, d1)
    # the tangential coordinate is unchanged
    assert abs(np.dot(x[0] - pts[0], ny)) < 1e-6 and abs(np.dot(x[1] - pts[1], nx)) < 1e-6


def test_capsule_side_and_cap():
    a, b, rad = np.array([-0.1, 0.0, 0.0]), np.array([0.1, 0.0, 0.0]), 0.03
    pts = np.array([[0.02, 0.035, 0.0], [0.13, 0.0, 0.01], [0.0, 0.0, -0.1]])
    x, _ = one_step(pts, lambda s: s.add_capsule(a, b, rad))
    side = math.hypot(x[0][1], x[0][2])
    cap = np.linalg.norm(x[1] - b)
    assert abs(side - (f32(rad) + R_P)) < TOL and abs(x[0][0] - 0.02) < 1e-7
    assert abs(cap - (f32(rad) + R_P)) < TOL
    assert np.allclose(x[2], pts[2])


def test_moving_sphere_leaves_particles_outside_its_final_position():
    # setSphere with a velocity sweeps the sphere through the substeps; after
    # the last substep (clock 0) no particle is inside the final sphere.
    c_end, rad, vel = np.array([0.0, 0.0, 0.0]), 0.05, np.array([3.0, 0.0, 0.0])
    pts = np.array([[0.055, 0.0, 0.0], [0.04, 0.03, 0.0], [0.0, 0.058, 0.0]])

    def add(s):
        s.add_prim(1, 2, [*c_end, rad], [0, 0, 0, 0.5], [*vel, 0.4])

    x, v = one_step(pts, add, substeps=4, iterations=2)
    dist = np.linalg.norm(x - c_end, axis=1)
    assert np.all(dist >= f32(rad) + R_P - TOL), dist
    assert x[0][0] > pts[0][0]  # pushed ahead of the moving sphere


def test_posed_capsule_follows_its_pose():
    cap = PosedPrimitive("capsule", [-0.1, 0.0, 0.0, 0.03], [0.1, 0.0, 0.0, 0.0])
    cap.pose([0.0, 0.0, 0.0], [0, 0, 0, 1])
    q = np.array([0, 0, math.sin(math.pi / 4), math.cos(math.pi / 4)])  # 90 degrees about z
    cap.pose([0.0, 0.2, 0.0], q, dt=1.0 / 60.0)
    assert np.allclose(cap.d0[:3], [0.0, 0.1, 0.0]) and np.allclose(cap.d1[:3], [0.0, 0.3, 0.0])
    assert np.allclose(cap.vel[:3], [0.0, 12.0, 0.0])
    assert np.allclose(cap.spin, [0.0, 0.0, (math.pi / 2) * 60.0])
