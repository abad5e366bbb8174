"""The oracle's local shape matching (shapeMatchImplicit.ts, rotation.ts)
pinned by the reference's own expectations
(tests/analytical/softbody/softbody-implicit.gpu.test.ts): a rigid
translation keeps the shape, a stiff spinning body keeps its angular
velocity, a pinned cantilever bends with >= 10 degrees between its ends, and
particles without edges keep the identity orientation."""
import math

import numpy as np
import pytest

from helpers import oracle_local_body, quat_to_mat, voxel_grid


@pytest.mark.parametrize("precision", ["f32", "f64"])
def test_rigid_translation_keeps_the_shape(precision):
    r = 0.05
    rest, edges = voxel_grid(3, 3, 3, r)
    n = len(rest)
    v = np.tile(np.float32([2.0, 0.0, 0.0]), (n, 1))
    s = oracle_local_body(rest, v, np.ones(n, np.float32), rest, edges, 1e-12, r, 4, 2, (0.0, 0.0, 0.0), precision)
    for _ in range(300):
        s.step(1 / 60)
    x = s.get()["x"]
    dev = np.linalg.norm((x - x.mean(0)) - rest, axis=1).max()
    if precision == "f64":
        assert dev < 1e-9, dev
        assert abs(x[:, 0].mean() - 2.0 * 300 / 60) < 1e-9
    else:
        # FP32 floor: the body sits ~10 m from the origin, where one ulp is
        # 1e-6 m and every substep's prediction rounds (the reference's own
        # FP32 run quotes 1e-3 m for the same test)
        assert dev < 2e-3, dev


@pytest.mark.parametrize("precision", ["f32", "f64"])
def test_spinning_body_keeps_its_angular_velocity(precision):
    r = 0.05
    rest, edges = voxel_grid(3, 3, 3, r)
    n = len(rest)
    om = 2.0
    v = np.stack([-om * rest[:, 1], om * rest[:, 0], np.zeros(n)], 1).astype(np.float32)
    s = oracle_local_body(rest, v, np.ones(n, np.float32), rest, edges, 1e-12, r, 4, 2, (0.0, 0.0, 0.0), precision)
    frames = 30
    for _ in range(frames):
        s.step(1 / 60)
    x = s.get()["x"]
    turned = math.atan2(x[14, 1] - x[13, 1], x[14, 0] - x[13, 0])
    expected = om * frames / 60
    assert 0.9 * expected < turned < 1.05 * expected, turned


def test_cantilever_bends_between_its_ends():
    r = 0.05
    nx, ny, nz = 8, 2, 2
    rest, edges = voxel_grid(nx, ny, nz, r)
    n = len(rest)
    fixed = np.abs(rest[:, 0] - rest[:, 0].min()) < 1e-6
    free = np.abs(rest[:, 0] - rest[:, 0].max()) < 1e-6
    w = np.where(fixed, 0.0, 1.0).astype(np.float32)
    s = oracle_local_body(rest, None, w, rest, edges, 1e-6, r, 8, 2, (0.0, -30.0, 0.0), "f32")
    for _ in range(60):
        s.step(1 / 60)
    q, _, _ = s.orientation()
    ang = lambda i: math.acos(max(-1.0, min(1.0, (np.trace(quat_to_mat(q[i])) - 1) / 2)))
    div = (np.mean([ang(i) for i in np.nonzero(free)[0]]) - np.mean([ang(i) for i in np.nonzero(fixed)[0]])) * 180 / math.pi
    assert abs(div) > 10, div
    assert np.all(np.isfinite(s.get()["x"]))


def test_particles_without_edges_keep_the_identity():
    r = 0.05
    rest, _ = voxel_grid(3, 3, 3, r)
    n = len(rest)
    s = oracle_local_body(rest, None, np.ones(n, np.float32), rest, np.zeros(0, np.uint32), 1e-6, r, 4, 2,
                          (0.0, -9.81, 0.0), "f32")
    for _ in range(20):
        s.step(1 / 60)
    q, _, _ = s.orientation()
    assert np.all(np.isfinite(q))
    err = max(np.abs(quat_to_mat(q[i]) - np.eye(3)).max() for i in range(n))
    assert err < 1e-4, err
