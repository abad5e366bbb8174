"""The oracle's surface field (SurfaceField.ts:101-369) pinned by closed forms:
one particle's phi is its distance minus the radius (the re-based Gaussian
mean is exact for a single sample), a particle on a voxel centre blurs into
the separable product of the normalised Gaussian taps, the volume fraction
sums to (spacing / voxel)^3 per particle, the wetted-wall proximity of a plane
is exp(-d / width) and nothing inside a collider is liquid; FP32 and FP64
agree on a stepped dam break."""
import math

import numpy as np
import pytest

import oracle as orc
from helpers import oracle_dam_break
from particle_fluids import scenes

R = 0.01
SPACING = 2 * R


def single(pos, precision="f32", **kw):
    s = orc.Scene(len(pos), R, precision)
    s.set_particles(np.asarray(pos, np.float32), kw.pop("v", None))
    return s


def voxel_centres(field, voxel, mn):
    nz, ny, nx, _ = field.shape
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    return np.stack([mn[0] + (x + 0.5) * voxel, mn[1] + (y + 0.5) * voxel, mn[2] + (z + 0.5) * voxel], -1)


def gauss_taps(voxel):
    sigma = 0.75 * SPACING / voxel
    taps = int(min(6, max(2, math.ceil(2.5 * sigma))))
    w = np.exp(-np.arange(-taps, taps + 1, dtype=np.float64) ** 2 / (2 * sigma * sigma))
    return taps, w / w.sum()


@pytest.mark.parametrize("precision", ["f32", "f64"])
def test_single_particle_phi_is_distance_minus_radius(precision):
    mn, mx = (-0.1, -0.1, -0.1), (0.1, 0.1, 0.1)
    p = np.array([[0.0123, -0.0071, 0.0049]])
    s = single(p, precision)
    field, voxel = s.surface_field(mn, mx, 40 ** 3, SPACING, (0, 1))
    c = voxel_centres(field, voxel, mn)
    d = np.linalg.norm(c - p.astype(np.float32).astype(np.float64), axis=-1) / voxel - 1.25 * R / voxel
    has = field[..., 1] > 1e-5 * ((SPACING / voxel) ** 3)
    assert has.sum() > 100
    err = np.abs(field[..., 0].astype(np.float64) - np.minimum(d, 6.0))[has].max()
    # 16-bit fixed point of the cloud-in-cell weights bounds the error
    assert err < 5e-4, err
    assert np.all(field[~has][..., 0] == 6.0)
    assert np.all(field[..., 3] == 1.0) and np.all(field[..., 2] == 0.0)


def test_voxel_centred_particle_blurs_into_separable_gaussian():
    mn, mx = (0.0, 0.0, 0.0), (0.4, 0.4, 0.4)
    s0 = single(np.zeros((1, 3)), "f64")
    _, voxel = s0.surface_field(mn, mx, 32 ** 3, SPACING, (0, 1))
    c = np.array([15, 17, 14])
    p = np.asarray(mn) + (c + 0.5) * voxel
    field, voxel2 = single(p[None], "f64").surface_field(mn, mx, 32 ** 3, SPACING, (0, 1))
    assert voxel2 == voxel
    taps, w = gauss_taps(voxel)
    vs = (SPACING / voxel) ** 3
    got = field[c[2] - taps:c[2] + taps + 1, c[1] - taps:c[1] + taps + 1, c[0] - taps:c[0] + taps + 1, 1]
    want = np.einsum("k,j,i->kji", w, w, w) * vs
    np.testing.assert_allclose(got, want, rtol=1e-9, atol=1e-12)
    # nothing reaches beyond the taps
    assert abs(field[..., 1].sum() - want.sum()) < 1e-9 * vs


@pytest.mark.parametrize("stretch", [0.0, 0.05])
def test_volume_fraction_sums_to_particle_volume(stretch):
    rng = np.random.default_rng(3)
    n = 300
    p = rng.uniform(-0.05, 0.05, (n, 3))
    v = rng.normal(0, 1.0, (n, 3)).astype(np.float32)
    mn, mx = (-0.15, -0.15, -0.15), (0.15, 0.15, 0.15)
    s = single(p, "f64", v=v)
    field, voxel = s.surface_field(mn, mx, 50 ** 3, SPACING, (0, n), stretch=stretch)
    vs = (SPACING / voxel) ** 3
    assert field[..., 1].max() < 4.0
    # each of 8 corners (times 3 samples) rounds to 1/65536 of a particle
    tol = vs * n * 8 * (3 if stretch else 1) * 0.5 / 65536
    assert abs(field[..., 1].sum() - n * vs) <= tol, (field[..., 1].sum(), n * vs)


def test_motion_stretch_trails_the_velocity():
    """Three samples on the segment p -> p - v * stretch (clamped to three
    spacings): the volume's centroid moves back by half the trail."""
    p = np.array([[0.0, 0.0, 0.0]])
    mn, mx = (-0.15, -0.1, -0.1), (0.1, 0.1, 0.1)
    for v, stretch, trail in (((2.0, 0.0, 0.0), 0.01, 0.02), ((50.0, 0.0, 0.0), 0.01, 3 * SPACING)):
        s = single(p, "f64", v=np.array([v], np.float32))
        field, voxel = s.surface_field(mn, mx, 64 ** 3, SPACING, (0, 1), stretch=stretch)
        c = voxel_centres(field, voxel, mn)
        w = field[..., 1]
        centroid = (c * w[..., None]).reshape(-1, 3).sum(0) / w.sum()
        assert abs(centroid[0] + trail / 2) < 1e-4 * voxel, (centroid, trail)
        assert abs(centroid[1]) < 1e-6 and abs(centroid[2]) < 1e-6


def test_wetted_plane_proximity_and_carving():
    mn, mx = (-0.1, -0.05, -0.1), (0.1, 0.15, 0.1)
    n = 50
    rng = np.random.default_rng(5)
    p = rng.uniform((-0.05, 0.0, -0.05), (0.05, 0.03, 0.05), (n, 3))
    s = single(p, "f64")
    s.set_loop(1, 1, (0.0, -9.81, 0.0), 4 * R, 64)
    s.add_plane((0, 1, 0), (0, 0, 0))
    width = 2.5 * R
    field, voxel = s.surface_field(mn, mx, 40 ** 3, SPACING, (0, n), wet=True)
    c = voxel_centres(field, voxel, mn)
    d = c[..., 1]
    np.testing.assert_allclose(field[..., 2], np.exp(-np.maximum(d, 0.0) / width), rtol=1e-12, atol=0)
    # no liquid inside the wall: phi >= depth below the plane
    assert np.all(field[..., 0] >= -d / voxel - 1e-12)
    below = d < -voxel
    assert below.any() and np.all(field[below][..., 0] > 0)
    # dry, distant voxels keep the band value lowered by the meniscus only within reach
    lift = 1.6 * R / voxel
    far = (field[..., 1] == 0) & (d > 0)
    reach = np.clip(1 - 6.0 / max(lift * 1.5, 1e-5), 0, 1)
    np.testing.assert_allclose(field[far][..., 0], 6.0 - lift * field[far][..., 2] * reach * reach, rtol=1e-12)
    # a dry field with the same walls is unwetted
    field0, _ = s.surface_field(mn, mx, 40 ** 3, SPACING, (0, n), wet=False)
    assert np.all(field0[..., 2] == 0.0)


def test_solids_raise_wall_proximity():
    mn, mx = (-0.1, -0.1, -0.1), (0.1, 0.1, 0.1)
    p = np.array([[0.0, 0.0, 0.0], [0.03, 0.0, 0.0]])
    s = single(p, "f64")
    field, voxel = s.surface_field(mn, mx, 40 ** 3, SPACING, (0, 1), solid_range=(1, 1))
    vs = (SPACING / voxel) ** 3
    taps, w = gauss_taps(voxel)
    # the solid's blurred mass times 1.5 volume scale, clamped to 1, is the proximity
    assert field[..., 2].max() == pytest.approx(min(1.0, field[..., 2].max()))
    assert field[..., 2].max() > 0
    # and solids add no liquid volume
    assert abs(field[..., 1].sum() - vs) < 8 * 0.5 / 65536 * vs
    # proximity sums to 1.5 vs per solid particle where unclamped
    if field[..., 2].max() < 1.0:
        assert abs(field[..., 2].sum() - 1.5 * vs) < 8 * 1.5 * vs / 65536


def test_fp32_field_agrees_with_fp64_on_dam_break():
    pos, r = scenes.dam_break_layout(2000)
    pos = scenes.jitter(pos, 2 * r)
    s32 = oracle_dam_break(pos, r, 2, 3, "f32")
    for _ in range(8):
        s32.step(1 / 60)
    st = s32.get()
    mn, mx = scenes.CONTAINER[:3], scenes.CONTAINER[3:]
    out = []
    for prec in ("f32", "f64"):
        s = orc.Scene(len(pos), r, prec)
        s.set_particles(st["x"].astype(np.float32), st["v"].astype(np.float32))
        out.append(s.surface_field(mn, mx, 200000, 2 * r, (0, len(pos)), stretch=0.02))
    (f32, v32), (f64, v64) = out
    assert v32 == v64
    near = np.abs(f64[..., 0]) < 5.0
    assert near.sum() > 1000
    assert np.abs(f32[..., 0] - f64[..., 0])[near].max() < 1e-3
    assert np.abs(f32[..., 1] - f64[..., 1]).max() < 1e-5
