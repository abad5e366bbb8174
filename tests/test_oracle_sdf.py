"""The oracle's SDF collider (SDFCollider.ts, sdf.ts:107-146) pinned by known
answers: three's truncating half-float conversion, exact trilinear
reconstruction of a linear field, a sphere's distance and normal, and a
particle dropped on an SDF floor coming to rest on it."""
import math

import numpy as np

import oracle as orc


def test_half_conversion_known_answers():
    cases = [
        (1.0, 0x3C00), (-2.0, 0xC000), (65504.0, 0x7BFF), (70000.0, 0x7BFF), (-1e9, 0xFBFF),
        (0.0, 0x0000), (-0.0, 0x8000), (1e-8, 0x0000), (0.1, 0x2E66), (2.0 ** -20, 0x0010),
        (2.0 ** -24, 0x0001), (0.5, 0x3800),
        # truncation, not rounding: 1 + 3/4 ulp stays at 1.0 (numpy rounds up)
        (1.0 + 3 * 2.0 ** -12, 0x3C00),
        # largest value below 1 + 1 ulp keeps the lower half
        (float(np.nextafter(np.float32(1.0 + 2.0 ** -10), np.float32(0))), 0x3C00),
    ]
    for v, bits in cases:
        assert orc.float_to_half(v) == bits, (v, hex(orc.float_to_half(v)), hex(bits))
    assert np.float16(1.0 + 3 * 2.0 ** -12).view(np.uint16) == 0x3C01  # the rounding the reference does not do
    for h in (0x3C00, 0x2E66, 0x0001, 0x7BFF, 0xC000):
        assert orc.half_to_float(h) == float(np.uint16(h).view(np.float16))


def field_scene(data, res, origin, voxel, **kw):
    s = orc.Scene(1, 0.01, kw.pop("precision", "f64"))
    s.set_particles(np.zeros((1, 3), np.float32))
    s.set_loop(1, 1, (0.0, 0.0, 0.0), 0.04, 64)
    s.add_sdf(data, res, origin, voxel, **kw)
    return s


def test_trilinear_reconstructs_a_linear_field():
    # phi = a.x + b sampled at voxel centres, values exact in half floats;
    # trilinear filtering and central differences reproduce it inside
    res = (8, 9, 10)
    origin = np.array([-0.5, -0.5625, -0.625])
    voxel = np.array([0.125, 0.125, 0.125])
    a = np.array([0.25, -0.5, 0.125])
    idx = np.stack(np.meshgrid(*[np.arange(n) for n in res], indexing="ij"), -1)
    centres = origin + (idx + 0.5) * voxel
    phi = (centres @ a + 0.0625).transpose(2, 1, 0)  # z, y, x order: x fastest
    for v in phi.reshape(-1)[:50]:
        assert orc.half_to_float(orc.float_to_half(v)) == np.float32(v)
    s = field_scene(phi, res, origin, voxel)
    rng = np.random.default_rng(1)
    for x in rng.uniform(-0.25, 0.25, (40, 3)):
        p, g = s.sdf_sample(0, x)
        assert abs(p - (x @ a + 0.0625)) < 1e-6, (x, p)
        assert np.allclose(g, a, atol=1e-6), g


def sphere_grid(radius, res, voxel):
    n = np.array(res)
    origin = -0.5 * n * voxel
    idx = np.stack(np.meshgrid(*[np.arange(k) for k in res], indexing="ij"), -1)
    c = origin + (idx + 0.5) * voxel
    return (np.linalg.norm(c, axis=-1) - radius).transpose(2, 1, 0), origin


def test_sphere_distance_and_normal_with_placement():
    res, voxel = (40, 40, 40), 0.01
    data, origin = sphere_grid(0.1, res, voxel)
    ang = 0.7
    rot = np.array([[math.cos(ang), 0, math.sin(ang)], [0, 1, 0], [-math.sin(ang), 0, math.cos(ang)]])
    pos, scale = np.array([0.3, 0.2, -0.1]), 1.5
    s = field_scene(data, res, origin, (voxel,) * 3, position=pos, rotation=rot, scale=scale)
    rng = np.random.default_rng(2)
    for _ in range(30):
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        x = pos + d * scale * rng.uniform(0.06, 0.16)
        p, g = s.sdf_sample(0, x)
        assert abs(p - (np.linalg.norm(x - pos) - 0.1 * scale)) < 0.01 * scale
        assert np.dot(g / np.linalg.norm(g), d) > 0.995


def test_particle_rests_on_an_sdf_floor():
    # a floor slab phi = y (field of a half space) under gravity: the
    # particle settles at phi = radius
    res = (8, 16, 8)
    origin = np.array([-0.4, -0.4, -0.4])
    voxel = np.array([0.1, 0.05, 0.1])
    idx = np.stack(np.meshgrid(*[np.arange(k) for k in res], indexing="ij"), -1)
    c = origin + (idx + 0.5) * voxel
    data = c[..., 1].transpose(2, 1, 0)
    r = 0.01
    s = orc.Scene(1, r, "f64")
    s.set_particles(np.array([[0.05, 0.1, 0.02]], np.float32))
    s.set_loop(4, 2, (0.0, -9.81, 0.0), 4 * r, 64)
    s.add_sdf(data, res, origin, voxel)
    for _ in range(120):
        s.step(1 / 60)
    st = s.get()
    assert abs(st["x"][0, 1] - r) < 1e-3, st["x"]
    assert np.abs(st["v"][0]).max() < 1e-2
