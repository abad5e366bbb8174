"""GPU SDFCollider (SDFCollider.ts, sdf.ts:107-146) against the FP32 oracle:
a fluid block dropped into a rotated, scaled spherical bowl given as an SDF."""
import math

import numpy as np
import pytest

import oracle as orc
import particle_fluids as pf
from helpers import rel_err
from particle_fluids import scenes

pytestmark = pytest.mark.gpu

TOL = 1e-5


def bowl_sdf(R=0.3, res=(24, 24, 24)):
    """phi = R - |x|: free space inside the ball, the collider outside."""
    n = np.array(res)
    voxel = 2.4 * R / n
    origin = -0.5 * n * voxel
    idx = np.stack(np.meshgrid(*[np.arange(k) for k in res], indexing="ij"), -1)
    c = origin + (idx + 0.5) * voxel
    data = (R - np.linalg.norm(c, axis=-1)).transpose(2, 1, 0).astype(np.float32)
    return dict(data=data, resolution=res, origin=origin, voxelSize=voxel)


def rotation_rows(q):
    """Rotation matrix rows as the library forms them (Quaternion.normalize,
    Matrix4.makeRotationFromQuaternion), in double."""
    q = np.asarray(q, np.float64)
    inv = 1.0 / math.sqrt(float((q * q).sum()))
    x, y, z, w = (float(v) * inv for v in q)
    x2, y2, z2 = x + x, y + y, z + z
    xx, xy, xz, yy, yz, zz = x * x2, x * y2, x * z2, y * y2, y * z2, z * z2
    wx, wy, wz = w * x2, w * y2, w * z2
    return np.array([[1 - (yy + zz), xy - wz, xz + wy], [xy + wz, 1 - (xx + zz), yz - wx],
                     [xz - wy, yz + wx, 1 - (xx + yy)]])


@pytest.mark.parametrize("fluid", [True, False])
def test_sdf_bowl_matches_fp32_oracle(fluid):
    print(f"[sdf] bowl fluid={fluid}", flush=True)
    r = 0.008
    sdf = bowl_sdf()
    q = (0.2, 0.4, -0.1, 0.9)
    pos_c, scale, thick = (0.1, 0.45, -0.05), 1.3, 0.002
    g = np.arange(-0.12, 0.12, 2 * r)
    pts = np.array([(x, y, z) for x in g for y in g for z in g], np.float32) + np.array(pos_c, np.float32)
    n = len(pts)
    ps = pf.ParticleSystem(n, r)
    ps.uploadParticles(pts, invMass=np.full(n, 1.0, np.float32))
    col = pf.SDFCollider(ps, sdf, position=pos_c, rotation=q, scale=scale, thickness=thick, muS=0.3, muK=0.2)
    mats = []
    if fluid:
        fl = pf.FluidSystem(ps, **scenes.FLUID_DEFAULTS)
        mats.append(fl)
    loop = pf.SimLoop(ps, substeps=2, iterations=3, materials=mats, colliders=[col])
    o = orc.Scene(n, r, "f32")
    o.set_particles(pts, w=np.full(n, 1.0, np.float32))
    h = 4 * r
    H = 1
    while H < 2 * n:
        H <<= 1
    o.set_loop(2, 3, (0.0, -9.81, 0.0), h, H)
    if fluid:
        o.add_fluid(0, n, 1000.0, 2 * r, h, 1e-4, viscosity=0.01, vorticity=0.02, surface_tension=0.1, adhesion=0.1)
    o.add_sdf(sdf["data"], sdf["resolution"], sdf["origin"], sdf["voxelSize"], position=pos_c, rotation=rotation_rows(q),
              scale=scale, thickness=thick, mus=0.3, muk=0.2)
    for step in range(12):
        loop.step(1 / 60)
        o.step(1 / 60)
    snap = ps.readback()
    ref = o.get()
    err = dict(x=rel_err(snap.positions[:, :3], ref["x"]), v=rel_err(snap.velocities[:, :3], ref["v"]))
    print(f"[sdf] rel err vs FP32 oracle after 12 steps {err}", flush=True)
    for k, e in err.items():
        assert e <= TOL, (k, e)
    # the particles stay inside the scaled bowl
    d = np.linalg.norm(snap.positions[:, :3] - np.array(pos_c), axis=1)
    assert d.max() <= 0.3 * scale - r + 5e-3, d.max()
