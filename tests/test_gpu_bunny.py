"""The demo's bunny dam break (demo/presets/dambreak.ts:9-66) at test size,
built from the reference's own bunny.sdf.bin: a column of water dropped into
a tank floods the Stanford bunny, an SDFCollider at the demo's yaw and scale.
The GPU state matches the FP32 oracle every step; the wetted surface field
over the demo's bounds (colliders = tank and bunny, motionStretch 0.02)
matches the oracle's over the same particles, and no liquid is drawn inside
the bunny."""
import math
import os

import numpy as np
import pytest

import oracle as orc
import particle_fluids as pf
from helpers import rel_err
from particle_fluids import formats
from test_gpu_sdf import rotation_rows

pytestmark = pytest.mark.gpu

HONEY = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "honey", "bunny.sdf.bin")
HALF_X, HALF_Z, COLUMN, LIFT, BUNNY_SCALE, BUNNY_YAW = 1.0, 0.45, 0.6, 0.2, 1.6, 2.16
HEIGHT = 0.5
R = 0.02
DT = 1.0 / 60.0
TANK = [((0, 1, 0), (0, 0, 0)), ((1, 0, 0), (-HALF_X, 0, 0)), ((-1, 0, 0), (HALF_X, 0, 0)),
        ((0, 0, 1), (0, 0, -HALF_Z)), ((0, 0, -1), (0, 0, HALF_Z))]


def column():
    s = 2 * R
    xs = np.arange(-HALF_X + R, -HALF_X + COLUMN - R + 1e-9, s)
    ys = np.arange(LIFT + R, LIFT + HEIGHT - R + 1e-9, s)
    zs = np.arange(-HALF_Z + R, HALF_Z - R + 1e-9, s)
    return np.array([(x, y, z) for z in zs for y in ys for x in xs], np.float32)


def test_bunny_dam_break_and_its_surface_match_oracle():
    sdf = formats.load_sdf(HONEY)
    pos = column()
    n = len(pos)
    print(f"[bunny] {n} particles, sdf {sdf['resolution']}", flush=True)
    yaw = BUNNY_YAW - math.pi / 2
    q = (0.0, math.sin(yaw / 2), 0.0, math.cos(yaw / 2))
    ps = pf.ParticleSystem(n, R)
    ps.uploadParticles(pos)
    fluid = pf.FluidSystem(ps, viscosity=0.01, surfaceTension=0.1, vorticity=0.02)
    walls = pf.PrimitiveSet(ps)
    for nrm, pt in TANK:
        walls.addPlane(nrm, pt)
    bunny = pf.SDFCollider(ps, sdf, rotation=q, scale=BUNNY_SCALE, muS=0.3, muK=0.2)
    loop = pf.SimLoop(ps, substeps=4, iterations=2, materials=[fluid], colliders=[walls, bunny])
    info = loop.info()
    o = orc.Scene(n, R, "f32")
    o.set_particles(pos)
    o.set_loop(4, 2, (0.0, -9.81, 0.0), info["cellSize"], info["hashTableSize"])
    o.add_fluid(0, n, 1000.0, 2 * R, 4 * R, 1e-4, viscosity=0.01, vorticity=0.02, surface_tension=0.1)
    for nrm, pt in TANK:
        o.add_plane(nrm, pt)
    o.add_sdf(sdf["data"], sdf["resolution"], sdf["origin"], sdf["voxelSize"], rotation=rotation_rows(q),
              scale=BUNNY_SCALE, mus=0.3, muk=0.2)
    errs = []
    for _ in range(45):
        loop.step(DT)
        o.step(DT)
        snap, ref = ps.readback(), o.get()
        errs.append((rel_err(snap.positions[:, :3], ref["x"]), rel_err(snap.velocities[:, :3], ref["v"])))
    print(f"[bunny] rel err (x, v) max {np.max(errs, axis=0)}", flush=True)
    assert np.max(errs) <= 1e-5, errs
    # the flood reached the bunny: some particles are within a few radii of its surface
    x = snap.positions[:, :3].astype(np.float64)
    assert x[:, 0].max() > -0.35

    mn = (-HALF_X - 0.03, -0.02, -HALF_Z - 0.03)
    mx = (HALF_X + 0.03, LIFT + 1.1, HALF_Z + 0.03)
    field = pf.SurfaceField(fluid, (mn, mx), motionStretch=0.02, wet=True)
    field.update()
    got = field.readback()
    of = orc.Scene(n, R, "f32")
    of.set_particles(snap.positions[:, :3], snap.velocities[:, :3])
    of.set_loop(4, 2, (0.0, -9.81, 0.0), info["cellSize"], info["hashTableSize"])
    for nrm, pt in TANK:
        of.add_plane(nrm, pt)
    of.add_sdf(sdf["data"], sdf["resolution"], sdf["origin"], sdf["voxelSize"], rotation=rotation_rows(q),
               scale=BUNNY_SCALE, mus=0.3, muk=0.2)
    want, voxel = of.surface_field(mn, mx, 600000, 2 * R, (0, n), stretch=0.02, wet=True)
    assert field.voxel == voxel and got.shape == want.shape
    err = np.abs(got[..., [0, 2]] - want[..., [0, 2]]).max()
    print(f"[bunny] field {got.shape}, max |phi, proximity - oracle| {err}", flush=True)
    assert err <= 2e-5
    assert np.array_equal(got[..., 1], want[..., 1])
    assert (got[..., 0] < 0).sum() > 500
    # voxel centres inside the bunny hold no liquid (phi > 0 there)
    nz, ny, nx, _ = got.shape
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    c = np.stack([mn[0] + (i + 0.5) * voxel, mn[1] + (j + 0.5) * voxel, mn[2] + (k + 0.5) * voxel], -1)
    rot = rotation_rows(q)
    local = (c @ rot) / BUNNY_SCALE  # rows^T applied: world -> local
    g = (local - np.asarray(sdf["origin"])) / np.asarray(sdf["voxelSize"]) - 0.5
    idx = np.floor(g).astype(int)
    ok = np.all((idx >= 0) & (idx < np.asarray(sdf["resolution"]) - 1), axis=-1)
    phi_s = sdf["data"].reshape(sdf["resolution"][2], sdf["resolution"][1], sdf["resolution"][0])
    deep = np.zeros(ok.shape, bool)
    sel = np.argwhere(ok)
    vals = phi_s[idx[ok][:, 2], idx[ok][:, 1], idx[ok][:, 0]]
    deep[tuple(sel.T)] = vals * BUNNY_SCALE < -3 * voxel
    assert deep.sum() > 100
    assert np.all(got[deep][..., 0] > 0)
    field.dispose()
