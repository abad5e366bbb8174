"""GPU local shape matching (shapeMatchImplicit.ts, rotation.ts) against the
FP32 oracle, plus the reference's behavioural checks on the GPU result."""
import math

import numpy as np
import pytest

import particle_fluids as pf
from helpers import oracle_local_body, quat_to_mat, rel_err, voxel_grid

pytestmark = pytest.mark.gpu


def gpu_local_body(pos, vel, w, rest, edges, compliance, r, substeps, iterations, gravity, colliders=()):
    n = len(pos)
    ps = pf.ParticleSystem(n, r)
    ps.uploadParticles(pos, velocities=vel, invMass=w)
    sb = pf.SoftbodySystem(ps, [dict(range=(0, n), restPositions=rest, compliance=compliance, edges=edges)],
                           shapeMatching="local")
    loop = pf.SimLoop(ps, substeps=substeps, iterations=iterations, gravity=gravity, materials=[sb],
                      colliders=list(colliders))
    return ps, sb, loop


@pytest.mark.parametrize("case", ["spin", "cantilever"])
def test_local_shape_matching_matches_fp32_oracle(case):
    print(f"[local] {case}", flush=True)
    r = 0.05
    if case == "spin":
        rest, edges = voxel_grid(3, 3, 3, r)
        n = len(rest)
        vel = np.stack([-2.0 * rest[:, 1], 2.0 * rest[:, 0], np.zeros(n)], 1).astype(np.float32)
        w = np.ones(n, np.float32)
        args = (1e-12, r, 4, 2, (0.0, 0.0, 0.0))
        steps = 30
    else:
        rest, edges = voxel_grid(8, 2, 2, r)
        n = len(rest)
        vel = np.zeros((n, 3), np.float32)
        w = np.where(np.abs(rest[:, 0] - rest[:, 0].min()) < 1e-6, 0.0, 1.0).astype(np.float32)
        args = (1e-6, r, 8, 2, (0.0, -30.0, 0.0))
        steps = 60
    ps, sb, loop = gpu_local_body(rest, vel, w, rest, edges, *args)
    o = oracle_local_body(rest, vel, w, rest, edges, *args, precision="f32")
    errs = []
    for k in range(steps):
        loop.step(1 / 60)
        o.step(1 / 60)
        if k in (0, 4, steps - 1):
            snap = ps.readback()
            ref = o.get()
            q, _, _ = o.orientation()
            e = dict(x=rel_err(snap.positions[:, :3], ref["x"]), v=rel_err(snap.velocities[:, :3], ref["v"]),
                     q=rel_err(snap.rotation, q))
            errs.append((k + 1, e))
    print(f"[local] rel err vs FP32 oracle {errs}", flush=True)
    # the first step has no spin yet, so no sine/cosine: bit-identical
    assert errs[0][1]["x"] <= 1e-6 and errs[0][1]["q"] <= 1e-6, errs[0]
    # later steps differ only by the device and host sin/cos/atan2 rounding
    for k, e in errs:
        assert e["x"] <= 1e-4 and e["q"] <= 1e-3, (k, e)
    snap = ps.readback()
    if case == "spin":
        x = snap.positions[:, :3]
        turned = math.atan2(x[14, 1] - x[13, 1], x[14, 0] - x[13, 0])
        assert 0.9 * 1.0 < turned < 1.05 * 1.0, turned
    else:
        ang = lambda i: math.acos(max(-1.0, min(1.0, (np.trace(quat_to_mat(snap.rotation[i])) - 1) / 2)))
        fixed = np.nonzero(np.abs(rest[:, 0] - rest[:, 0].min()) < 1e-6)[0]
        free = np.nonzero(np.abs(rest[:, 0] - rest[:, 0].max()) < 1e-6)[0]
        div = (np.mean([ang(i) for i in free]) - np.mean([ang(i) for i in fixed])) * 180 / math.pi
        print(f"[local] cantilever divergence {div:.2f} deg", flush=True)
        assert abs(div) > 10, div


def test_local_body_frames_follow_the_body():
    # bodyCenters / bodyRotations stay current under local matching
    r = 0.05
    rest, edges = voxel_grid(3, 3, 3, r)
    n = len(rest)
    vel = np.tile(np.float32([0.5, 0.0, 0.0]), (n, 1))
    ps, sb, loop = gpu_local_body(rest, vel, np.ones(n, np.float32), rest, edges, 1e-9, r, 4, 2, (0.0, 0.0, 0.0))
    for _ in range(30):
        loop.step(1 / 60)
    c, R = sb.readbackFrames()
    x = ps.readback().positions[:, :3].astype(np.float64)
    assert np.allclose(c[0], x.mean(0), atol=1e-4), (c[0], x.mean(0))
    assert abs(c[0][0] - 0.25) < 1e-2
    assert np.abs(R[0] - np.eye(3)).max() < 1e-3
