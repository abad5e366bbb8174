"""ViscositySolver in the oracle (src/fluids/ViscositySolver.ts:78-123).

Pinned by what the solve must satisfy rather than by its formula: nu = 0
changes nothing; a uniform velocity field is a fixed point; and after enough
Jacobi sweeps the velocities solve the implicit system the sweeps iterate,
v_i (1 + rate W_i) - rate sum_j w_ij v_j = v0_i with w_ij = poly6(r_ij^2) V,
W_i = sum_j w_ij (self included) and rate = nu dt 10 / h^2
(ViscositySolver.ts:96-104), checked against a brute-force neighbour set."""
import numpy as np
import pytest

import oracle as orc
from helpers import oracle_dam_break
from particle_fluids import scenes

R = 0.01
H = 4 * R


def sparse_block(nx, ny, nz, gap=2.6):
    g = np.stack(np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij"), -1).reshape(-1, 3)
    return g * gap * R


def sparse_scene(pos, vel, precision, nu=None, iters=64, substeps=1):
    """A block below rest density (no pressure correction), no gravity, no
    colliders, no other fluid effects: only the viscosity solve changes v."""
    n = len(pos)
    s = orc.Scene(n, R, precision)
    s.set_particles(pos, vel)
    s.set_loop(substeps, 1, (0.0, 0.0, 0.0), H, 1024)
    s.add_fluid(0, n, 1000.0, 2 * R, H, 1e-4)
    if nu is not None:
        s.add_viscosity(0, nu, iters)
    return s


@pytest.mark.parametrize("precision", ["f32", "f64"])
def test_zero_viscosity_changes_nothing(precision):
    pos, r = scenes.dam_break_layout(2000)
    pos = scenes.jitter(pos, 2 * r)
    a = oracle_dam_break(pos, r, 2, 2, precision)
    b = oracle_dam_break(pos, r, 2, 2, precision)
    b.add_viscosity(0, 0.0, 5)
    for _ in range(3):
        a.step(1 / 60)
        b.step(1 / 60)
    sa, sb = a.get(), b.get()
    assert np.array_equal(sa["x"], sb["x"]) and np.array_equal(sa["v"], sb["v"])


def test_uniform_velocity_is_a_fixed_point():
    pos = sparse_block(8, 8, 8)
    u = np.array([0.3, -0.2, 0.1])
    s = sparse_scene(pos, np.tile(u, (len(pos), 1)), "f64", nu=5.0, iters=16)
    s.step(1 / 60)
    v = s.get()["v"]
    u32 = u.astype(np.float32).astype(np.float64)  # particle state enters through f32 storage
    assert np.abs(v - u32).max() < 1e-12, np.abs(v - u32).max()


def test_converged_sweeps_solve_the_implicit_system():
    pos = sparse_block(7, 6, 5)
    rng = np.random.default_rng(3)
    vel = rng.normal(0, 0.2, pos.shape)
    nu, dt = 0.02, 1 / 60
    plain = sparse_scene(pos, vel, "f64")
    solved = sparse_scene(pos, vel, "f64", nu=nu, iters=64)
    plain.step(dt)
    solved.step(dt)
    a, b = plain.get(), solved.get()
    x = b["x"]
    assert np.allclose(a["x"], x, atol=0, rtol=0)  # the solve touches velocities only
    v0, v = a["v"], b["v"]
    spacing, h = 2 * R, H
    poly6 = 315.0 / (64.0 * np.pi * h ** 9)
    d2 = ((x[:, None, :] - x[None, :, :]) ** 2).sum(-1)
    w = np.where(d2 < h * h, np.maximum(h * h - d2, 0.0) ** 3 * poly6, 0.0) * spacing ** 3
    rate = nu * dt * 10.0 / (h * h)
    res = v * (1.0 + rate * w.sum(1))[:, None] - rate * (w @ v) - v0
    assert np.abs(v - v0).max() > 1e-3          # the solve did something
    assert np.abs(res).max() < 1e-10 * np.abs(v0).max(), np.abs(res).max()
