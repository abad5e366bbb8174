"""ViscositySolver on the GPU (src/fluids/ViscositySolver.ts) against the
FP32 oracle, in a batch against each scene alone, and through
Simulation.addFluid(thickness=...)."""
import math

import numpy as np
import pytest

import particle_fluids as pf
from particle_fluids import scenes
from helpers import oracle_dam_break, rel_err

pytestmark = pytest.mark.gpu

DT = 1.0 / 60.0


def thick_dam_break(budget, nu, iters, seed=0xC0FFEE):
    pos, r = scenes.dam_break_layout(budget)
    pos = scenes.jitter(pos, 2 * r, seed)
    ps = pf.ParticleSystem(len(pos), r)
    ps.uploadParticles(pos)
    fluid = pf.FluidSystem(ps, **scenes.FLUID_DEFAULTS)
    mats = [fluid]
    solver = None
    if nu is not None:
        solver = pf.ViscositySolver(fluid, nu, iterations=iters)
        mats.append(solver)
    walls = pf.PrimitiveSet(ps)
    mn, mx = scenes.CONTAINER[:3], scenes.CONTAINER[3:]
    for n, which in scenes.PLANES:
        walls.addPlane(n, mn if which == "min" else mx)
    loop = pf.SimLoop(ps, substeps=2, iterations=4, materials=mats, colliders=[walls])
    return dict(ps=ps, fluid=fluid, solver=solver, walls=walls, loop=loop, positions=pos, radius=r)


def test_viscosity_solver_matches_fp32_oracle():
    print("[viscosity] honey-like dam break vs oracle", flush=True)
    sc = thick_dam_break(3000, 20.0, 12)
    o = oracle_dam_break(sc["positions"], sc["radius"], 2, 4, "f32", H=sc["loop"].info()["hashTableSize"])
    o.add_viscosity(0, 20.0, 12)
    errs = []
    for _ in range(4):
        sc["loop"].step(DT)
        o.step(DT)
        s, ref = sc["ps"].readback(), o.get()
        errs.append((rel_err(s.positions[:, :3], ref["x"]), rel_err(s.velocities[:, :3], ref["v"])))
    print(f"[viscosity] rel err per step (x, v) {errs}", flush=True)
    for ex, ev in errs:
        assert ex <= 1e-5 and ev <= 1e-4, errs


def test_batch_with_different_sweeps_matches_alone():
    spec = [(3000, 20.0, 9), (5000, 5.0, 12), (4000, None, 0)]
    alone = [thick_dam_break(b, nu, it, seed=b) for b, nu, it in spec]
    mixed = [thick_dam_break(b, nu, it, seed=b) for b, nu, it in spec]
    batch = pf.Batch([m["loop"] for m in mixed])
    batch.useGraph(True)
    for _ in range(3):
        for a in alone:
            a["loop"].step(DT)
        batch.step(DT)
    for i, (a, m) in enumerate(zip(alone, mixed)):
        ra, rm = a["ps"].readback(), m["ps"].readback()
        assert np.array_equal(ra.positions, rm.positions) and np.array_equal(ra.velocities, rm.velocities), i


def test_simulation_thickness_damps_the_flow():
    def run(thickness):
        sim = pf.Simulation(particles=3000, container=(-0.5, 0, -0.3, 0.5, 0.8, 0.3), substeps=2, iterations=4)
        sim.addFluid((-0.5, 0, -0.3, -0.1, 0.5, 0.3), thickness=thickness)
        for _ in range(30):
            sim.step(DT)
        v = sim.particleSystem.readback().velocities[:, :3].astype(np.float64)
        n = sim.particleCount
        sim.dispose()
        return 0.5 * (v * v).sum(1).mean(), n
    ke_water, n = run(None)
    ke_honey, _ = run(20.0)
    print(f"[viscosity] mean KE water {ke_water:.4e} honey {ke_honey:.4e} ({n} particles, "
          f"{math.floor(16 * float(np.cbrt(n / 1e4)) + 0.5)} sweeps)", flush=True)
    assert ke_honey < ke_water


def test_solver_must_follow_its_fluid():
    sc = thick_dam_break(2000, None, 0)
    ps = pf.ParticleSystem(100, 0.01)
    ps.uploadParticles(np.random.default_rng(0).random((100, 3)) * 0.1)
    f = pf.FluidSystem(ps)
    sb_solver = pf.ViscositySolver(f, 1.0)
    with pytest.raises(pf.PFError, match="directly after its FluidSystem"):
        pf.SimLoop(ps, materials=[sb_solver, f])
    with pytest.raises(pf.PFError, match="iterations must be an integer from 1 to 64"):
        pf.ViscositySolver(f, 1.0, iterations=65)
    with pytest.raises(pf.PFError, match="viscosity must be"):
        pf.ViscositySolver(f, -1.0)
    scenes.dispose_scene(sc)
