"""GPU fluid boundaries (Akinci volumes, adhesion, boundary reaction) against
the FP32 oracle, and momentum conservation of the reaction on the GPU."""
import numpy as np
import pytest

import particle_fluids as pf
from helpers import boundary_layout, oracle_boundary_scene, rel_err
from particle_fluids import scenes

pytestmark = pytest.mark.gpu

TOL = 1e-5


def gpu_boundary_scene(pos, w, nf, nb, radius, dynamic, substeps=2, iterations=4, gravity=(0.0, -9.81, 0.0),
                       effects=None, planes=True):
    ps = pf.ParticleSystem(len(pos), radius)
    ps.uploadParticles(pos, invMass=w)
    fx = dict(scenes.FLUID_DEFAULTS) if effects is None else effects
    fluid = pf.FluidSystem(ps, range=(0, nf), **fx)
    fluid.addBoundary((nf, nb), dynamic=dynamic)
    cols = []
    if planes:
        walls = pf.PrimitiveSet(ps)
        mn, mx = scenes.CONTAINER[:3], scenes.CONTAINER[3:]
        for nrm, which in scenes.PLANES:
            walls.addPlane(nrm, mn if which == "min" else mx)
        cols.append(walls)
    loop = pf.SimLoop(ps, substeps=substeps, iterations=iterations, gravity=gravity, materials=[fluid],
                      colliders=cols)
    return dict(ps=ps, fluid=fluid, loop=loop)


@pytest.mark.parametrize("dynamic,plate_w", [(False, 0.0), (True, 0.0), (True, 1.0 / (0.5 * 1000.0 * 0.02 ** 3))])
def test_boundary_scene_matches_fp32_oracle(dynamic, plate_w):
    print(f"[boundary] dynamic={dynamic} plate_w={plate_w:.3g}", flush=True)
    r = 0.01
    pos, w, nf, nb = boundary_layout(radius=r, plate_w=plate_w)
    g = gpu_boundary_scene(pos, w, nf, nb, r, dynamic)
    o = oracle_boundary_scene(pos, w, nf, nb, r, dynamic)
    for step in range(3):
        g["loop"].step(1 / 60)
        o.step(1 / 60)
        vol_g = g["ps"].readbackBoundaryVolume()[nf:]
        vol_o = o.boundary_volumes()[nf:]
        assert np.all(vol_o > 0)
        e_vol = rel_err(vol_g, vol_o)
        snap = g["ps"].readback()
        ref = o.get()
        err = dict(x=rel_err(snap.positions[:, :3], ref["x"]), v=rel_err(snap.velocities[:, :3], ref["v"]),
                   rho=rel_err(g["fluid"].readbackDensity()[:nf], ref["rho"][:nf]), vol=e_vol)
        print(f"[boundary] step {step + 1} rel err vs FP32 oracle {err}", flush=True)
        for k, e in err.items():
            assert e <= TOL, (step, k, e)


def test_gpu_reaction_conserves_momentum():
    r = 0.01
    pos, w, nf, nb = boundary_layout(radius=r)
    pos[:nf, 1] -= 0.035
    m_fluid = 1000.0 * (2 * r) ** 3
    w[nf:] = 1.0 / (0.5 * m_fluid)
    g = gpu_boundary_scene(pos, w, nf, nb, r, True, substeps=1, iterations=4, gravity=(0.0, 0.0, 0.0),
                           effects={}, planes=False)
    g["loop"].step(1 / 120)
    v = g["ps"].readback().velocities[:, :3].astype(np.float64)
    m = np.concatenate([np.full(nf, m_fluid), np.full(nb, 0.5 * m_fluid)])
    p = (m[:, None] * v).sum(0)
    scale = (m[:, None] * np.abs(v)).sum()
    print(f"[boundary] momentum {p} scale {scale}", flush=True)
    assert scale > 0
    assert np.abs(p).max() <= 1e-5 * scale
