"""Closed-form and brute-force checks of the oracle's fluid passes.

Each check builds a tiny scene, steps the oracle and compares one pass's
output with a value computed here from the reference's formulas by another
route: an all-pairs brute force instead of the grid walk, a closed form on a
symmetric input, or the two-particle case solved by hand. The file:line of
the reference code each check follows is given in its docstring.

Every check takes the oracle library to run (None: oracle/liboracle.so), so
tests/test_oracle_planted_mistakes.py can run the same checks against builds
of the oracle with one mistake planted, and show that each mistake fails at
least one of them.
"""
import math

import numpy as np

import oracle as orc


# kernels.ts:22-47: coefficients in FP64 on the host.
def poly6_coef(h):
    return 315.0 / (64.0 * math.pi * h ** 9)


def spiky_coef(h):
    return 45.0 / (math.pi * h ** 6)


def W(r2, h):
    """Poly6 from r^2 (kernels.ts:64-69)."""
    return poly6_coef(h) * max(h * h - r2, 0.0) ** 3


def grad_w(r, h):
    """Spiky gradient at offset r = x_i - x_j (kernels.ts:89-107)."""
    d = float(np.sqrt(np.dot(r, r)))
    if d == 0.0:
        return np.zeros(3)
    return -spiky_coef(h) * max(h - d, 0.0) ** 2 * np.asarray(r, np.float64) / d


def next_pow2(v):
    p = 1
    while p < v:
        p <<= 1
    return p


def fluid_scene(pos, h, spacing, library=None, prec="f64", vel=None, w=None, iterations=1, substeps=1,
                gravity=(0.0, 0.0, 0.0), compliance=1e-4, count=None, **effects):
    """One fluid over the first `count` particles (all by default), grid cell
    h and table nextPow2(2n), as SimLoop sizes them (loop.ts:151-162)."""
    pos = np.asarray(pos, np.float32)
    n = len(pos)
    s = orc.Scene(n, spacing / 2, prec, library=library)
    s.set_particles(pos, None if vel is None else np.asarray(vel, np.float32), w)
    s.set_loop(substeps, iterations, gravity, h, next_pow2(2 * n))
    s.add_fluid(0, n if count is None else count, 1000.0, spacing, h, compliance, **effects)
    return s


def lattice(k, spacing, origin=(0.0, 0.0, 0.0)):
    """k^3 points, x fastest (the reference tests' loop order)."""
    g = np.stack(np.meshgrid(np.arange(k), np.arange(k), np.arange(k), indexing="ij"), -1).reshape(-1, 3)[:, ::-1]
    return (np.asarray(origin) + g * spacing).astype(np.float32)


def centred_lattice(k, spacing):
    """(2k+1)^3 points at m * spacing, m in [-k, k], x fastest."""
    m = np.arange(-k, k + 1)
    g = np.stack(np.meshgrid(m, m, m, indexing="ij"), -1).reshape(-1, 3)[:, ::-1]
    return (g * spacing).astype(np.float32), g


# ---------------------------------------------------------------------------
# neighbour list (neighborList.ts:45-72 over query.ts:24-70)
# ---------------------------------------------------------------------------
def check_neighbour_sets_5_lattice(library=None):
    """pair-list-correctness.gpu.test.ts:23-90: on a 5^3 lattice (spacing
    0.025, h 0.04) every list holds exactly the brute-force set within h, self
    included, and the overflow flag stays down."""
    pos = lattice(5, 0.025)
    h = 0.04
    s = fluid_scene(pos, h, 0.025, library, prec="f32", iterations=0)
    s.step(1.0 / 60.0)
    nl, cnt = s.neighbors()
    p = pos.astype(np.float64)
    d2 = ((p[:, None, :] - p[None, :, :]) ** 2).sum(-1)
    for i in range(len(pos)):
        want = set(np.nonzero(d2[i] < h * h)[0].tolist())
        assert cnt[i] == len(want), f"particle {i}: {cnt[i]} neighbours, brute force {len(want)}"
        assert set(nl[i, :cnt[i]].tolist()) == want, f"particle {i}: neighbour set differs"
    assert not s.flags()["neighbors"]


def check_neighbour_filter_is_strict(library=None):
    """neighborList.ts:55-60 keeps |x_i - x_j|^2 < h^2: a particle exactly h
    away (dyadic coordinates, so the distance is exact) is not a neighbour."""
    h = 0.5
    pos, g = centred_lattice(1, 0.5)
    s = fluid_scene(pos, h, 0.25, library, prec="f32", iterations=0)
    s.step(1.0 / 60.0)
    nl, cnt = s.neighbors()
    c = int(np.nonzero((g == 0).all(1))[0][0])
    assert cnt[c] == 1 and nl[c, 0] == c, "only the particle itself is closer than h"


def check_neighbour_cap_and_order(library=None):
    """pair-list-overflow.gpu.test.ts:21-74: MAX_NEIGHBORS + 32 particles all
    within h: every list holds 64 entries and the flag is raised. They share
    one cell, so the walk visits one bucket in ascending particle order and
    each list keeps particles 0..63 (the stable walk order; sort.ts:187-192
    leaves it to the build)."""
    n = 64 + 32
    side = math.ceil(n ** (1.0 / 3.0))
    pos = lattice(side, 0.05)[:n]
    p = pos.astype(np.float64)
    assert (((p[:, None] - p[None]) ** 2).sum(-1) < 1.0).all()
    s = fluid_scene(pos, 1.0, 0.05, library, prec="f32", iterations=0)
    s.step(1.0 / 60.0)
    nl, cnt = s.neighbors()
    assert (cnt == 64).all()
    assert (nl[:, :64] == np.arange(64, dtype=np.uint32)[None, :]).all()
    assert s.flags()["neighbors"]


def check_neighbour_walk_order_brute_force(library=None):
    """Walk order and truncation (query.ts:37-69, neighborList.ts:55-66):
    stencil buckets dz -> dy -> dx around the query's cell, particles of a
    bucket in ascending order, the first 64 within h kept. The expected lists
    come from an all-pairs FP32 distance test sorted by (stencil position of
    the neighbour's cell, index), on a jittered lattice compressed so that
    some lists overflow."""
    rng = np.random.default_rng(0x51ED)
    sp = 0.02
    h = 2.0 * sp
    pos = lattice(9, 0.75 * sp, (0.1, 0.2, 0.3))
    pos = (pos + rng.uniform(-0.2, 0.2, pos.shape) * sp).astype(np.float32)
    s = fluid_scene(pos, h, sp, library, prec="f32", iterations=0)
    s.step(1.0 / 60.0)
    nl, cnt = s.neighbors()
    cell = np.floor(pos / np.float32(h)).astype(np.int64)  # IEEE FP32 division, origin 0
    h2 = np.float32(h * h)
    overflowed = 0
    for i in range(len(pos)):
        d = pos[i][None, :] - pos  # FP32, as the walk subtracts
        r2 = (d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2]
        js = np.nonzero(r2 < h2)[0]
        off = cell[js] - cell[i]
        assert (np.abs(off) <= 1).all()
        stencil = (off[:, 2] + 1) * 9 + (off[:, 1] + 1) * 3 + (off[:, 0] + 1)
        want = js[np.lexsort((js, stencil))][:64]
        overflowed += len(js) > 64
        assert cnt[i] == len(want), f"particle {i}: count {cnt[i]}, expected {len(want)}"
        assert (nl[i, :cnt[i]] == want).all(), f"particle {i}: list order differs"
    assert overflowed > 0, "the input must exercise the 64-entry cap"
    assert s.flags()["neighbors"]


# ---------------------------------------------------------------------------
# density and lambda (lambda.ts:30-71)
# ---------------------------------------------------------------------------
def check_density_5_lattice(library=None):
    """density-pair-list-cpu-parity.gpu.test.ts:33-149: the density of every
    particle of a 5^3 lattice (spacing 0.025, h 0.04, m = rho0 spacing^3)
    equals the brute-force SPH sum; FP32 within the reference test's
    envelope (5e-3 absolute, 1e-5 relative), FP64 within 1e-12."""
    pos = lattice(5, 0.025)
    h, sp = 0.04, 0.025
    p = pos.astype(np.float64)
    d2 = ((p[:, None] - p[None]) ** 2).sum(-1)
    m = 1.0 / float(np.float32(1.0 / (1000.0 * sp ** 3)))  # 1 / invMass, invMass stored as f32
    poly = poly6_coef(h)
    want = (np.where(d2 < h * h, m * np.maximum(h * h - d2, 0.0) ** 3 * poly, 0.0)).sum(1)
    for prec, abs_tol, rel_tol in (("f32", 5e-3, 1e-5), ("f64", None, 1e-12)):
        s = fluid_scene(pos, h, sp, library, prec=prec, iterations=1)
        s.step(1.0 / 240.0)
        rho = s.get()["rho"]
        err = np.abs(rho - want)
        if abs_tol is not None:
            assert err.max() < abs_tol
        assert (err / np.maximum(want, 1.0)).max() < rel_tol, f"{prec}: density differs from the SPH sum"


def two_particle_case(library, direction=(1.0, 2.0, 2.0), sep=0.04, h=0.08, dt=1.0 / 60.0, compliance=1e-4):
    """Two fluid particles `sep` apart with particle spacing h (so the pair is
    compressed, C > 0), one iteration, no gravity: returns the oracle's
    state and fields with the closed-form lambda and delta x worked out from
    lambda.ts:40-66 and positionDelta.ts:44-72."""
    u = np.asarray(direction, np.float64)
    u /= np.linalg.norm(u)
    pos = np.array([[0.3, 0.3, 0.3], 0.3 + sep * u], np.float32)
    s = fluid_scene(pos, h, h, library, prec="f64", iterations=1, compliance=compliance)
    s.step(dt)
    rho0 = 1000.0
    mass = rho0 * h ** 3
    w = float(np.float32(1.0 / mass))  # invMass is stored as f32 (FluidSystem.ts:118)
    m = 1.0 / w                        # emitNeighborMass: 1 / invMass (shared.ts:52-62)
    p = pos.astype(np.float64)
    r = p[0] - p[1]
    r2 = float(r @ r)
    rho = m * (W(0.0, h) + W(r2, h))
    C = rho / rho0 - 1.0
    g = m * grad_w(r, h)               # the self term's gradient is zero
    denom = ((g @ g) * w + (g @ g) * w) / rho0 ** 2 + compliance / dt ** 2
    lam = -C / max(denom, 1e-20)
    dx0 = (w / rho0) * (lam + lam) * m * grad_w(r, h)
    return s, dict(rho=rho, C=C, lam=lam, dx=np.stack([dx0, -dx0]), pos=p)


def check_two_particle_lambda_and_delta(library=None):
    """A compressed pair: both particles get lambda = -C / (|grad C|^2 terms +
    alpha/dt^2) with the inverse mass of each side (lambda.ts:60-66), and
    equal and opposite corrections (lambda_i + lambda_j) m grad W w_i / rho0
    along the line between them (positionDelta.ts:44-72), pushing them apart.
    The mass (0.512 kg) is not 1, so each inverse mass factor shows."""
    s, want = two_particle_case(library)
    assert want["C"] > 0.1
    st, f = s.get(), s.fluid_fields()
    assert np.allclose(st["rho"], want["rho"], rtol=1e-12, atol=0)
    assert np.allclose(f["lam"], want["lam"], rtol=1e-11, atol=0), f"lambda {f['lam']} vs {want['lam']}"
    assert np.allclose(f["dx"], want["dx"], rtol=1e-10, atol=1e-18)
    sep0 = np.linalg.norm(want["pos"][0] - want["pos"][1])
    sep1 = np.linalg.norm(st["x"][0] - st["x"][1])
    assert sep1 > sep0
    assert np.abs(st["x"] - (want["pos"] + f["dx"])).max() < 1e-15


def check_uncompressed_pair_has_zero_lambda(library=None):
    """lambda.ts:59-63: C <= 0 gives lambda 0 (unilateral), and the density is
    written anyway: two particles at the default spacing h/2, 0.9h apart."""
    h = 0.1
    pos = np.array([[0.3, 0.3, 0.3], [0.3 + 0.9 * h, 0.3, 0.3]], np.float32)
    s = fluid_scene(pos, h, h / 2, library, prec="f64", iterations=1)
    s.step(1.0 / 60.0)
    m = 1.0 / float(np.float32(1.0 / (1000.0 * (h / 2) ** 3)))
    d2 = float(((pos[0].astype(np.float64) - pos[1]) ** 2).sum())
    want = m * (W(0.0, h) + W(d2, h))
    assert want < 1000.0
    assert np.allclose(s.get()["rho"], want, rtol=1e-12)
    f = s.fluid_fields()
    assert (f["lam"] == 0).all() and (f["dx"] == 0).all()


def check_symmetric_lattice_delta(library=None):
    """A compressed cubic lattice (spacing 7/256 at h = 1/16, mass at h/2),
    centred on the origin with exact coordinates: lambda is symmetric and
    delta x antisymmetric under point reflection, the centre does not move,
    and particles on the axes move outward (the block expands)."""
    h = 0.0625
    pos, g = centred_lattice(3, 7.0 / 256.0)
    s = fluid_scene(pos, h, h / 2, library, prec="f64", iterations=1)
    s.step(1.0 / 60.0)
    f = s.fluid_fields()
    nl, cnt = s.neighbors()
    assert cnt.max() < 64, "no list may be cut, or the symmetry breaks"
    k = 3
    idx = lambda m: (m[..., 2] + k) * 49 + (m[..., 1] + k) * 7 + (m[..., 0] + k)  # noqa: E731
    mirror = idx(-g)
    lam, dx = f["lam"], f["dx"]
    assert (lam < 0).any()
    scale = np.abs(dx).max()
    assert scale > 0
    assert np.abs(lam - lam[mirror]).max() <= 1e-12 * np.abs(lam).max()
    assert np.abs(dx + dx[mirror]).max() <= 1e-12 * scale
    c = idx(np.zeros(3, np.int64))
    assert np.abs(dx[c]).max() <= 1e-12 * scale
    for a in range(3):
        m = np.zeros(3, np.int64)
        for step in (1, 2, 3):
            m[a] = step
            assert dx[idx(m)][a] > 0, f"axis {a}, +{step}: correction must point outward"
            assert abs(dx[idx(m)][(a + 1) % 3]) <= 1e-12 * scale


# ---------------------------------------------------------------------------
# XSPH and vorticity (velocity.ts:36-119), ViscositySolver
# ---------------------------------------------------------------------------
def check_xsph_pair_and_uniform_drift(library=None):
    """velocity.ts:59-66: dv_i = c sum_j V_j (v_j - v_i) W. A uniform drift
    gives exactly zero; for a moving particle beside a resting one the change
    slows the mover and drags the other along, equal and opposite."""
    h, sp = 0.04, 0.02
    pos = lattice(4, sp, (0.1, 0.1, 0.1))
    vel = np.tile(np.array([0.3, -0.2, 0.5], np.float32), (len(pos), 1))
    s = fluid_scene(pos, h, sp, library, vel=vel, iterations=0, viscosity=0.1, vorticity=0.5)
    s.step(1.0 / 60.0)
    f = s.fluid_fields()
    # advect's (x* - x) / dt leaves rounding of order 1e-15 m/s in the drift
    assert np.abs(f["dv"]).max() < 1e-13 and np.abs(f["omega"]).max() < 1e-11
    assert np.abs(s.get()["v"] - vel.astype(np.float64)).max() < 1e-12

    c = 0.25
    pos = np.array([[0.1, 0.1, 0.1], [0.1 + 0.6 * h, 0.1, 0.1]], np.float32)
    vel = np.array([[0.0, 1.0, 0.0], [0.0, 0.0, 0.0]], np.float32)
    dt = 1e-4
    s = fluid_scene(pos, h, sp, library, vel=vel, iterations=0, viscosity=c)
    s.step(dt)
    st, f = s.get(), s.fluid_fields()
    x1 = st["x"]
    Wr = W(float(((x1[0] - x1[1]) ** 2).sum()), h)
    v = vel.astype(np.float64)
    want0 = c * sp ** 3 * (v[1] - v[0]) * Wr
    assert np.allclose(f["dv"][0], want0, rtol=1e-9, atol=1e-15)
    assert np.allclose(f["dv"][1], -want0, rtol=1e-9, atol=1e-15)
    assert f["dv"][0][1] < 0 < f["dv"][1][1]
    assert np.allclose(st["v"], v + f["dv"], rtol=0, atol=1e-10)


def check_viscous_diffusion(library=None):
    """viscous-diffusion.gpu.test.ts:11-58: a 3x3x3 block with shear
    vx in {-1, 0, 1} across x and uniform drift vy = 0.5 (h 0.08, spacing
    0.04, one substep of one iteration, no gravity); the ViscositySolver
    (viscosity 3, 16 sweeps) after the step halves the shear energy or
    better, keeps vy at 0.5 to 3 decimals, and stays bounded."""
    i = np.arange(27)
    pos = np.stack([(i % 3) * 0.04, (i // 3 % 3) * 0.04 + 1.0, (i // 9) * 0.04], 1).astype(np.float32)
    vel = np.stack([(i % 3) - 1.0, np.full(27, 0.5), np.zeros(27)], 1).astype(np.float32)
    runs = []
    for solver in (False, True):
        s = orc.Scene(27, 0.02, "f32", library=library)
        s.set_particles(pos, vel)
        s.set_loop(1, 1, (0.0, 0.0, 0.0), 0.08, next_pow2(54))
        s.add_fluid(0, 27, 1000.0, 0.04, 0.08, 1e-4)
        if solver:
            s.add_viscosity(0, 3.0, 16)
        s.step(1.0 / 60.0)
        runs.append(s.get()["v"])
    before, after = runs
    assert np.isfinite(after).all()
    assert (np.abs(after[:, 1] - 0.5) < 5e-4).all()
    assert (np.abs(after[:, 0]) <= 1.01).all()
    assert (after[:, 0] ** 2).sum() < 0.5 * (before[:, 0] ** 2).sum()


def check_rigid_rotation_vorticity(library=None):
    """velocity.ts:54: omega_i = sum_j V_j (v_i - v_j) x grad W. For a rigid
    rotation v = Omega x x on a cubic lattice the centre's omega is
    (2/3) Omega sum_j V_j f(r_j) r_j^2 with grad W = -f(r) r, parallel to
    Omega with a positive factor (a cubic neighbourhood's second moment is
    isotropic)."""
    h, sp = 0.04, 0.02
    pos, g = centred_lattice(3, sp)
    Om = np.array([0.0, 0.0, 2.0])
    vel = np.cross(Om[None, :], pos.astype(np.float64)).astype(np.float32)
    s = fluid_scene(pos, h, sp, library, vel=vel, iterations=0, vorticity=0.02)
    s.step(1e-5)
    f = s.fluid_fields()
    c = int(np.nonzero((g == 0).all(1))[0][0])
    p = pos.astype(np.float64)
    acc = 0.0
    for j in range(len(p)):
        r = p[c] - p[j]
        d = float(np.sqrt(r @ r))
        if 0 < d and d * d < h * h:
            acc += sp ** 3 * spiky_coef(h) * (h - d) ** 2 / d * d * d
    want = 2.0 / 3.0 * Om * acc
    om = f["omega"][c]
    assert om[2] > 0
    assert np.abs(om - want).max() <= 1e-6 * np.linalg.norm(want)
    assert abs(f["omlen"][c] - np.linalg.norm(om)) <= 1e-12 * np.linalg.norm(om)


def check_vorticity_gradient_and_confinement(library=None):
    """velocity.ts:74-119. The shear v = (0, a x^2, 0) has vorticity
    (0, 0, 2ax); on a lattice the SPH estimate at an interior particle is
    exactly linear in x, so eta = sum_j V (|w_j| - |w_i|) grad W points along
    +x (towards larger |omega|), and the confinement change
    eps dt (eta/|eta|) x omega is -eps dt |omega| along y."""
    h, sp = 0.04, 0.02
    eps, dt, a = 0.5, 1e-5, 1.0
    k = 11
    pos = lattice(k, sp, (0.1, -0.1, -0.1))
    x0 = pos.astype(np.float64)
    vel = np.zeros_like(pos)
    vel[:, 1] = (a * x0[:, 0] ** 2).astype(np.float32)
    s = fluid_scene(pos, h, sp, library, vel=vel, iterations=0, vorticity=eps)
    s.step(dt)
    st, f = s.get(), s.fluid_fields()
    m = np.stack(np.meshgrid(np.arange(k), np.arange(k), np.arange(k), indexing="ij"), -1).reshape(-1, 3)[:, ::-1]
    inner = ((m >= 4) & (m <= 6)).all(1)
    om, eta = f["omega"][inner], f["eta"][inner]
    assert (om[:, 2] > 0).all()
    assert np.abs(om[:, :2]).max() <= 1e-3 * np.abs(om[:, 2]).min()
    slope = om[:, 2] / x0[inner, 0]
    assert slope.max() - slope.min() <= 1e-3 * slope.mean()
    assert (eta[:, 0] > 0).all()
    assert np.abs(eta[:, 1:]).max() <= 1e-3 * eta[:, 0].min()
    v_adv = (st["x"] - x0) / dt
    dv = (st["v"] - v_adv)[inner]
    want = -eps * dt * np.linalg.norm(om, axis=1)
    assert (dv[:, 1] < 0).all()
    assert np.abs(dv[:, 1] - want).max() <= 2e-3 * np.abs(want).max()
    assert np.abs(dv[:, [0, 2]]).max() <= 2e-3 * np.abs(want).max()


# ---------------------------------------------------------------------------
# surface tension and adhesion (cohesion.ts:25-103, adhesion.ts:16-60)
# ---------------------------------------------------------------------------
def cohesion_spline(r, h):
    """Akinci et al. 2013, eq. 2 as cohesion.ts:46-55 writes it."""
    coef = 32.0 / (math.pi * h ** 9)
    if not (0.0 < r <= h):
        return 0.0
    if 2.0 * r > h:
        return coef * (h - r) ** 3 * r ** 3
    return coef * (2.0 * (h - r) ** 3 * r ** 3 - h ** 6 / 64.0)


def adhesion_spline(r, h):
    """Akinci et al. 2013, eq. 7 as adhesion.ts:16-27 writes it."""
    if not (2.0 * r > h and r <= h):
        return 0.0
    return 0.007 / h ** 3.25 * max(-4.0 * r * r / h + 6.0 * r - 2.0 * h, 0.0) ** 0.25


IMPULSE_SCALE = math.floor(2 ** 30 / 50)  # Accumulator(50): ticks per m/s (accumulator.ts:55-68)


def check_cohesion_pair(library=None):
    """cohesion.ts:63-103 on one pair at distances spanning the spline's
    branches (h/4, h/2, 3h/4, 0.95h): dv_i = (-gamma m^2 C(r) r_hat
    - gamma m (n_i - n_j)) K / m dt with n = h V grad W (cohesion.ts:25-39)
    and K = 1 (density is still rho0 before the first iteration,
    FluidSystem.ts:121-126), rounded to the impulse accumulator's ticks;
    the partner gets the exact negation."""
    h, sp, gamma, dt = 0.04, 0.02, 1.0, 1.0 / 60.0
    mass = 1000.0 * sp ** 3
    for frac in (0.25, 0.5, 0.75, 0.95):
        u = np.array([2.0, -1.0, 2.0]) / 3.0
        pos = np.array([[0.2, 0.2, 0.2], 0.2 + frac * h * u], np.float32)
        s = fluid_scene(pos, h, sp, library, iterations=0, surface_tension=gamma)
        s.step(dt)
        p = pos.astype(np.float64)
        off = p[0] - p[1]
        r = float(np.sqrt(off @ off))
        n_i = h * sp ** 3 * grad_w(off, h)
        curv = -(n_i - (-n_i)) * gamma * mass
        coh = -(off / r) * gamma * mass * mass * cohesion_spline(r, h)
        dv = (coh + curv) / mass * dt
        want = np.rint(dv * IMPULSE_SCALE) / IMPULSE_SCALE
        v = s.get()["v"]
        assert np.abs(v[0] - want).max() <= 1e-9, f"r = {frac} h: {v[0]} vs {want}"
        assert np.abs(v[0] + v[1]).max() <= 1e-13
        assert np.abs(want).max() > 1e3 / IMPULSE_SCALE


def check_cohesion_spline_knots(library=None):
    """The spline is continuous at h/2 (both branches give coef h^6/64) and
    vanishes at h; with the curvature term cancelled by symmetry (three
    particles in a row, the middle one's normal is zero) the outer pair's
    pull on the middle particle is zero and the middle particle does not
    move, whatever the spline."""
    h = 0.0625  # dyadic: the three positions and their distances are exact
    assert abs(cohesion_spline(h / 2, h) - 32.0 / (math.pi * h ** 9) * h ** 6 / 64.0) < 1e-9 * cohesion_spline(h / 2, h)
    assert cohesion_spline(h, h) == 0.0
    assert cohesion_spline(h / 4, h) < 0 < cohesion_spline(3 * h / 4, h)
    pos = np.array([[0.25 - 0.5 * h, 0.25, 0.25], [0.25, 0.25, 0.25], [0.25 + 0.5 * h, 0.25, 0.25]], np.float32)
    s = fluid_scene(pos, h, h / 2, library, iterations=0, surface_tension=1.0)
    s.step(1.0 / 60.0)
    v = s.get()["v"]
    assert np.abs(v[1]).max() <= 1e-12
    assert v[0][0] > 0 > v[2][0]


def check_adhesion_pair(library=None):
    """adhesion.ts:34-60: a fluid particle beside one static boundary particle
    (Akinci volume 1 / W(0) for a lone boundary particle, boundaryVolume.ts:26-52)
    gets dv = -beta psi A(r) r_hat dt with psi = rho0 V_b, rounded to ticks;
    zero at and below h/2."""
    h, sp, beta, dt = 0.0625, 0.03125, 2.0, 1.0 / 60.0  # dyadic: r = h/2 is exact
    Vb = 1.0 / W(0.0, h)
    for frac in (0.4, 0.5, 0.6, 0.75, 0.9, 0.999):
        pos = np.array([[0.25, 0.25, 0.25], [0.25 + frac * h, 0.25, 0.25]], np.float32)
        s = orc.Scene(2, sp / 2, "f64", library=library)
        s.set_particles(pos, None, np.zeros(2, np.float32))
        s.set_loop(1, 0, (0.0, 0.0, 0.0), h, next_pow2(4))
        s.add_fluid(0, 1, 1000.0, sp, h, 1e-4, adhesion=beta)
        s.add_boundary(0, 1, 1, dynamic=False)
        s.step(dt)
        assert abs(s.boundary_volumes()[1] - Vb) <= 1e-12 * Vb
        p = pos.astype(np.float64)
        off = p[0] - p[1]
        r = float(np.sqrt(off @ off))
        dv = -(off / r) * beta * 1000.0 * Vb * adhesion_spline(r, h) * dt
        want = np.rint(dv * IMPULSE_SCALE) / IMPULSE_SCALE
        v = s.get()["v"]
        assert np.abs(v[0] - want).max() <= 1e-9, f"r = {frac} h: {v[0]} vs {want}"
        assert (v[1] == 0).all()
        if frac > 0.5:
            assert want[0] > 1e3 / IMPULSE_SCALE, "the fluid particle is pulled towards the boundary"
        else:
            assert (want == 0).all()


CHECKS = [
    check_neighbour_sets_5_lattice,
    check_neighbour_filter_is_strict,
    check_neighbour_cap_and_order,
    check_neighbour_walk_order_brute_force,
    check_density_5_lattice,
    check_two_particle_lambda_and_delta,
    check_uncompressed_pair_has_zero_lambda,
    check_symmetric_lattice_delta,
    check_xsph_pair_and_uniform_drift,
    check_viscous_diffusion,
    check_rigid_rotation_vorticity,
    check_vorticity_gradient_and_confinement,
    check_cohesion_pair,
    check_cohesion_spline_knots,
    check_adhesion_pair,
]
