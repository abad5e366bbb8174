"""Sphere, box and capsule colliders (primitives.ts:85-217, PrimitiveSet.ts)
on the GPU against the FP32 oracle: static shapes under a falling block of
fluid, and shapes moved every step (setSphere with a velocity, a box and a
capsule posed by setPose, which sweep through the substeps)."""
import math

import numpy as np
import pytest

import oracle as orc
import particle_fluids as pf
from helpers import PosedPrimitive, rel_err

pytestmark = pytest.mark.gpu

DT = 1.0 / 60.0
R = 0.01


def block(nx, ny, nz, lo):
    g = np.stack(np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij"), -1).reshape(-1, 3)
    return (np.asarray(lo) + (g + 0.5) * 2 * R).astype(np.float32)


def build(pos, fluid):
    n = len(pos)
    ps = pf.ParticleSystem(n, R)
    ps.uploadParticles(pos)
    mats = []
    if fluid:
        mats.append(pf.FluidSystem(ps, viscosity=0.01, surfaceTension=0.1, vorticity=0.02, adhesion=0.1))
    else:
        ps.setInvMass((0, n), 1.0)
    walls = pf.PrimitiveSet(ps)
    H = 1
    while H < 2 * n:
        H <<= 1
    o = orc.Scene(n, R, "f32")
    o.set_particles(pos, None, None if fluid else np.ones(n, np.float32))
    return ps, mats, walls, o, min(H, 1 << 20)


def finish(ps, mats, walls, o, H, fluid, substeps=2, iterations=3):
    loop = pf.SimLoop(ps, substeps=substeps, iterations=iterations, materials=mats, colliders=[walls])
    cell = loop.info()["cellSize"]
    o.set_loop(substeps, iterations, (0.0, -9.81, 0.0), cell if fluid else 0.0, loop.info()["hashTableSize"])
    if fluid:
        o.add_fluid(0, ps.capacity, 1000.0, 2 * R, 4 * R, 1e-4, viscosity=0.01, vorticity=0.02,
                    surface_tension=0.1, adhesion=0.1)
    return loop


def compare(ps, o):
    s, ref = ps.readback(), o.get()
    return rel_err(s.positions[:, :3], ref["x"]), rel_err(s.velocities[:, :3], ref["v"])


@pytest.mark.parametrize("fluid", [True, False])
def test_static_shapes_match_oracle(fluid):
    print(f"[prims] static shapes, fluid={fluid}", flush=True)
    pos = block(12, 10, 12, (-0.12, 0.12, -0.12))
    ps, mats, walls, o, H = build(pos, fluid)
    q = (0.0, 0.0, math.sin(math.radians(20) / 2), math.cos(math.radians(20) / 2))
    walls.addPlane((0, 1, 0), (0, 0, 0))
    walls.addSphere((-0.05, 0.06, 0.0), 0.04)
    walls.addBox((0.07, 0.05, 0.02), (0.04, 0.02, 0.05), rotation=q, muS=0.3, muK=0.2)
    walls.addCapsule((-0.1, 0.02, -0.08), (0.1, 0.03, -0.08), 0.025)
    walls.addSphere((0.0, 0.2, 0.0), 0.5, invert=True)
    o.add_plane((0, 1, 0), (0, 0, 0))
    o.add_sphere((-0.05, 0.06, 0.0), 0.04)
    o.add_box((0.07, 0.05, 0.02), (0.04, 0.02, 0.05), q, mus=0.3, muk=0.2)
    o.add_capsule((-0.1, 0.02, -0.08), (0.1, 0.03, -0.08), 0.025)
    o.add_sphere((0.0, 0.2, 0.0), 0.5, invert=True)
    loop = finish(ps, mats, walls, o, H, fluid)
    errs = []
    for k in range(12):
        loop.step(DT)
        o.step(DT)
        errs.append(compare(ps, o))
    print(f"[prims] rel err (x, v) first {errs[0]} last {errs[-1]}", flush=True)
    for ex, ev in errs:
        assert ex <= 1e-5 and ev <= 1e-4, errs


def test_moving_shapes_match_oracle():
    print("[prims] moving shapes", flush=True)
    pos = block(10, 8, 10, (-0.1, 0.08, -0.1))
    ps, mats, walls, o, H = build(pos, True)
    walls.addPlane((0, 1, 0), (0, 0, 0))
    s_sphere = walls.addSphere((-0.2, 0.05, 0.0), 0.04)
    s_box = walls.addBox((0.0, 0.0, 0.0), (0.03, 0.02, 0.06))
    s_cap = walls.addCapsule((-0.05, 0.0, 0.0), (0.05, 0.0, 0.0), 0.02)
    o.add_plane((0, 1, 0), (0, 0, 0))
    o.add_sphere((-0.2, 0.05, 0.0), 0.04)
    o.add_box((0.0, 0.0, 0.0), (0.03, 0.02, 0.06))
    o.add_capsule((-0.05, 0.0, 0.0), (0.05, 0.0, 0.0), 0.02)
    box = PosedPrimitive("box", [0, 0, 0, 0], [0.03, 0.02, 0.06, 0])
    cap = PosedPrimitive("capsule", [-0.05, 0, 0, 0.02], [0.05, 0, 0, 0])
    loop = finish(ps, mats, walls, o, H, True)
    errs = []
    for k in range(12):
        t = k * DT
        c = (-0.2 + 1.5 * t, 0.05, 0.0)
        v = (1.5, 0.0, 0.0)
        walls.setSphere(s_sphere, c, 0.04, v)
        o.set_prim(s_sphere, 1, 2, [*c, 0.04], [0, 0, 0, 0.5], [*v, 0.4])
        a = 0.8 * t
        qb = (0.0, math.sin(a / 2), 0.0, math.cos(a / 2))
        pb = (0.08, 0.04 + 0.01 * math.sin(6 * t), 0.0)
        qc = (math.sin(a / 2), 0.0, 0.0, math.cos(a / 2))
        pc = (-0.02, 0.03, 0.05 - 0.5 * t)
        walls.setPose(s_box, pb, qb, DT if k else 0.0)
        walls.setPose(s_cap, pc, qc, DT if k else 0.0)
        box.pose(pb, qb, DT if k else 0.0)
        cap.pose(pc, qc, DT if k else 0.0)
        o.set_prim(s_box, *box.oracle_args())
        o.set_prim(s_cap, *cap.oracle_args())
        loop.step(DT)
        o.step(DT)
        errs.append(compare(ps, o))
    print(f"[prims] rel err (x, v) first {errs[0]} last {errs[-1]}", flush=True)
    for ex, ev in errs:
        assert ex <= 1e-5 and ev <= 1e-4, errs
