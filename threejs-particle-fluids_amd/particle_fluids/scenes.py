"""Workload scenes: the dam break and the isolated sort inputs.

Dam break (examples/fluid.ts geometry): container [-0.5,0.5]x[0,0.8]x[-0.3,0.3],
fluid box [-0.5,-0.1]x[0,0.5]x[-0.3,0.3], particles sized by Simulation's
budget fit, Simulation's default fluid effects (Simulation.ts:348-352) and
five container planes (Simulation.ts:777-787). The lattice gets a seeded
+-1 % spacing jitter from the reference tests' LCG
(tests/analytical/hashgrid/hashgrid-random-stress.gpu.test.ts:21-27).
"""
import math

import numpy as np

from . import FluidSystem, ParticleSystem, PrimitiveSet, SimLoop, SoftbodySystem, fill_box, fit_budget
from .simulation import _euler_xyz_quat, _quat_apply

CONTAINER = (-0.5, 0.0, -0.3, 0.5, 0.8, 0.3)
FLUID_BOX = (-0.5, 0.0, -0.3, -0.1, 0.5, 0.3)
FLUID_DEFAULTS = dict(viscosity=0.01, surfaceTension=0.1, vorticity=0.02, adhesion=0.1)
PLANES = (((0, 1, 0), "min"), ((1, 0, 0), "min"), ((-1, 0, 0), "max"), ((0, 0, 1), "min"), ((0, 0, -1), "max"))


def lcg_uniform(seed, n):
    """First n draws u = state / 2^32 of the LCG state = imul(state, 1664525) +
    1013904223 (mod 2^32), vectorised by composing the affine recurrence."""
    # x_{k+1} = a x_k + c mod 2^32; x_k = A_k x_0 + C_k with (A_k, C_k) composed by doubling.
    a, c, mod = 1664525, 1013904223, 1 << 32
    x0 = (seed & 0xFFFFFFFF) or 1
    out = np.empty(n, np.uint64)
    block = 4096
    # coefficients for steps 1..block
    A = np.empty(block, np.uint64)
    Cc = np.empty(block, np.uint64)
    ak, ck = 1, 0
    for k in range(block):
        ak, ck = (ak * a) % mod, (ck * a + c) % mod
        A[k], Cc[k] = ak, ck
    x = x0
    done = 0
    while done < n:
        m = min(block, n - done)
        vals = (A[:m] * np.uint64(x) + Cc[:m]) % np.uint64(mod)
        out[done:done + m] = vals
        x = int(vals[-1])
        done += m
    return out.astype(np.float64) / 4294967296.0


def dam_break_layout(budget):
    """(positions (N,3) float32, particle radius) of Simulation's fitted layout."""
    r, n = fit_budget(budget, CONTAINER, [FLUID_BOX])
    shrunk = [CONTAINER[i] + r for i in range(3)] + [CONTAINER[3 + i] - r for i in range(3)]
    lo = [max(FLUID_BOX[i], shrunk[i]) for i in range(3)]
    hi = [min(FLUID_BOX[3 + i], shrunk[3 + i]) for i in range(3)]
    pos = fill_box(lo, hi, 2 * r)
    assert len(pos) == n
    return pos, r


def jitter(positions, spacing, seed=0xC0FFEE, fraction=0.01):
    """Add a uniform +-fraction*spacing offset per axis, LCG draws in particle order."""
    u = lcg_uniform(seed, positions.size).reshape(positions.shape)
    return (positions.astype(np.float64) + (2.0 * u - 1.0) * fraction * spacing).astype(np.float32)


def build_dam_break(budget, substeps=2, iterations=4, seed=0xC0FFEE, context=None, use_graph=False):
    """The dam break scene on the GPU: dict of ps, fluid, walls, loop, positions, radius."""
    pos, r = dam_break_layout(budget)
    pos = jitter(pos, 2 * r, seed)
    ps = ParticleSystem(len(pos), r, context=context)
    ps.uploadParticles(pos)
    fluid = FluidSystem(ps, **FLUID_DEFAULTS)
    walls = PrimitiveSet(ps)
    mn, mx = CONTAINER[:3], CONTAINER[3:]
    for n, which in PLANES:
        walls.addPlane(n, mn if which == "min" else mx)
    loop = SimLoop(ps, substeps=substeps, iterations=iterations, materials=[fluid], colliders=[walls])
    if use_graph:
        loop.useGraph(True)
    return dict(ps=ps, fluid=fluid, walls=walls, loop=loop, positions=pos, radius=r)


# Soft-body cubes for the batched and served scene mixes.
CUBE_SIDE = 13          # 13^3 = 2,197 particles, the lattice cube nearest 2,000
CUBE_DENSITY = 500.0    # Simulation's soft-body density default (Simulation.ts:409, 1052)
CUBE_GAP = 0.1          # free space between neighbouring cubes, in cube widths
CUBE_DROP = 0.05        # initial height above the floor, in cube widths


def particle_cube(center, side, spacing):
    """A side^3 lattice cube centred on `center`, surface particles first
    (tests/perf/_scenes/_helpers.ts:88-111). Returns (positions, surfaceCount)."""
    i = np.arange(side)
    z, y, x = np.meshgrid(i, i, i, indexing="ij")  # x fastest, then y, then z
    idx = np.stack([x.ravel(), y.ravel(), z.ravel()], 1)
    surf = np.any((idx == 0) | (idx == side - 1), axis=1)
    off = (idx - (side - 1) / 2.0) * spacing + np.asarray(center, np.float64)
    pos = np.concatenate([off[surf], off[~surf]])
    return pos, int(surf.sum())


def softbody_compliance(softness, count):
    """Shape-matching compliance from softness (src/simulation/handles.ts:246-248)."""
    return 10.0 ** (-6.0 + softness * 3.0) * (count / 200.0)


def mixed_layout(fluid_budget, bodies, seed):
    """A dam break fitted to `fluid_budget` plus `bodies` soft cubes of
    CUBE_SIDE^3 particles at the fluid's particle spacing.

    The cubes rest in one layer on the floor of an extension of the tank past
    x = 0.5 (rows of cubes across z, as many rows as needed), so every cube
    fits at any particle radius; the dam-break water reaches them as it
    spreads. Stiffness k is uniform in [0.1, 1.0] per cube from `seed`;
    softness = 1 - k and compliance follows softbodyCompliance.
    Returns dict(positions, inv_mass, fluid_count, bodies, radius, container).
    """
    pos, r = dam_break_layout(fluid_budget)
    pos = jitter(pos, 2 * r, seed)
    nf = len(pos)
    container = list(CONTAINER)
    out = dict(fluid_count=nf, radius=r, bodies=[], container=tuple(container))
    if bodies == 0:
        out.update(positions=pos, inv_mass=None)
        return out
    spacing = 2 * r
    width = CUBE_SIDE * spacing
    pitch = (1.0 + CUBE_GAP) * width
    depth = CONTAINER[5] - CONTAINER[2]
    per_row = max(1, int(depth // pitch))
    rows = (bodies + per_row - 1) // per_row
    container[3] = CONTAINER[3] + rows * pitch
    rng = np.random.default_rng(seed)
    stiffness = rng.uniform(0.1, 1.0, bodies)
    cubes = []
    for b in range(bodies):
        row, col = divmod(b, per_row)
        c = (CONTAINER[3] + (row + 0.5) * pitch, width / 2 + CUBE_DROP * width,
             CONTAINER[2] + (col + 0.5) * depth / per_row)
        cp, surf = particle_cube(c, CUBE_SIDE, spacing)
        start = nf + sum(len(q) for q in cubes)
        out["bodies"].append(dict(range=(start, len(cp)), surfaceCount=surf,
                                  compliance=softbody_compliance(1.0 - stiffness[b], len(cp))))
        cubes.append(cp)
    allpos = np.concatenate([pos.astype(np.float64)] + cubes).astype(np.float32)
    w = np.empty(len(allpos), np.float32)
    w[:nf] = 0.0  # FluidSystem sets the fluid range's inverse mass
    w[nf:] = 1.0 / (CUBE_DENSITY * spacing ** 3)
    out.update(positions=allpos, inv_mass=w, container=tuple(container))
    return out


# Particle contacts Simulation turns on when a scene has soft bodies
# (Simulation.ts:821-836).
SIMULATION_CONTACT = {"muS": 0.3, "muK": 0.2}


def build_mixed_scene(fluid_budget, bodies, seed, substeps=2, iterations=4, context=None, use_graph=False,
                      evictable=False, contact=None, fluid_options=None):
    """One scene of the batched / served mixes on the GPU: the dam break with
    Simulation's default fluid effects (or `fluid_options`), `bodies` globally
    shape-matched cubes and the five planes of the (extended) tank.
    `evictable` places every device buffer in one arena so a Server can move
    the scene to host memory. `contact` is SimLoop's option; None turns
    contacts on as Simulation does: with SIMULATION_CONTACT whenever the scene
    has cubes."""
    lay = mixed_layout(fluid_budget, bodies, seed)
    return build_layout_scene(lay, substeps, iterations, context, use_graph, evictable, contact, fluid_options)


def build_layout_scene(lay, substeps=2, iterations=4, context=None, use_graph=False, evictable=False, contact=None,
                       fluid_options=None):
    """The GPU scene of a mixed_layout-shaped dict (positions, inv_mass,
    fluid_count, bodies, radius, container); see build_mixed_scene."""
    pos, nf, r = lay["positions"], lay["fluid_count"], lay["radius"]
    bodies = len(lay["bodies"])
    ps = ParticleSystem(len(pos), r, context=context, evictable=evictable)
    ps.uploadParticles(pos, invMass=lay["inv_mass"])
    fluid = FluidSystem(ps, range=(0, nf), **(FLUID_DEFAULTS if fluid_options is None else fluid_options))
    mats = [fluid]
    soft = None
    if bodies:
        soft = SoftbodySystem(ps, lay["bodies"])
        mats.append(soft)
    walls = PrimitiveSet(ps)
    mn, mx = lay["container"][:3], lay["container"][3:]
    for n, which in PLANES:
        walls.addPlane(n, mn if which == "min" else mx)
    if contact is None and bodies:
        # Only the cubes' particles search for contacts (the water keeps its own
        # particles apart), so 8 records per cube particle bound the pairs;
        # readbackOverflow().contacts reports it if they ever do not.
        soft_count = sum(b["range"][1] for b in lay["bodies"])
        contact = dict(SIMULATION_CONTACT, maxContacts=8 * soft_count)
    elif contact is None:
        contact = False
    loop = SimLoop(ps, substeps=substeps, iterations=iterations, materials=mats, colliders=[walls],
                   contact=contact)
    if use_graph:
        loop.useGraph(True)
    return dict(ps=ps, fluid=fluid, soft=soft, walls=walls, loop=loop, layout=lay)


# Elastic studies (demo/presets/elastic.ts:78-230): twenty baked soft forms
# dropped into a tank between two plates that squeeze them.
ELASTIC_NAMES = ("donut", "croissant", "banana", "ginger-bread")
ELASTIC_BODY_BUDGETS = (250, 500, 750, 1250, 2500)  # PARTICLE_LEVELS / 20 (demo/types.ts:35-49)
ELASTIC_BODIES = 20
ELASTIC_RELEASE_HEIGHT = 0.55


def scaled_substeps(base, particles, dimensions=3):
    """scaledSubsteps (demo/presets/shared.ts:61-64)."""
    scale = (particles / 10000.0) ** (1.0 / dimensions)
    return max(base, math.ceil(base * scale - 0.1))


def elastic_layout(templates, particles=5000, seed=0x9E3779B9):
    """Particle state of the elastic studies scene from baked templates (a
    dict name -> SampledBody from formats.load_sampled_body at the chosen
    per-body level): twenty forms, staggered, jittered and tumbled by the
    demo's seeded generator, each weighing the same at every level.
    Returns dict(positions, inv_mass, bodies, radius, per_body)."""
    from .formats import seeded_random
    per_body = min(ELASTIC_BODY_BUDGETS, key=lambda b: (abs(b - particles / ELASTIC_BODIES), b))
    radius = 0.015 * (200.0 / per_body) ** (1.0 / 3.0)
    random = seeded_random(seed)
    pos, inv, bodies = [], [], []
    start = 0
    for n in range(ELASTIC_BODIES):
        shape = templates[ELASTIC_NAMES[n % len(ELASTIC_NAMES)]]
        count = len(shape["positions"])
        layer = n // 12
        cx = ((n % 4) - 1.5) * 0.44 + (0.22 if layer else 0.0) + (random() - 0.5) * 0.12
        cy = ELASTIC_RELEASE_HEIGHT + layer * 0.55 + random() * 0.14
        cz = ((n // 4) % 3 - 1) * 0.47 + (random() - 0.5) * 0.12
        e = (random() * math.pi * 2, random() * math.pi * 2, random() * math.pi * 2)
        q = _euler_xyz_quat(e)
        p = _quat_apply(q, shape["positions"].astype(np.float64)) + np.array([cx, cy, cz])
        pos.append(p)
        inv.append(np.full(count, count / 20.0))
        bodies.append(dict(range=(start, count), surfaceCount=shape["surfaceCount"], edges=shape["edges"]))
        start += count
    return dict(positions=np.concatenate(pos).astype(np.float32), inv_mass=np.concatenate(inv).astype(np.float32),
                bodies=bodies, radius=radius, per_body=per_body)


def elastic_plate_distance(time, compression=0.8):
    """Where the squeezing plates stand at `time` (elastic.ts:205-211)."""
    cycle = (1.0 - math.cos(max(0.0, time - 1.0) * 1.05)) / 2.0
    return max(0.1, 1.06 - cycle * compression * 1.0875)


def build_elastic_scene(templates, particles=5000, softness=0.9, gravity=6.0, context=None, substeps=None,
                        iterations=2):
    """The elastic studies scene on the GPU: locally shape-matched forms with
    selfCollision, the tank (tank(1.05, 0.87), shared.ts:26-39), two plates
    and the demo's contacts (muS 0.12, muK 0.06, 20 records per particle).
    `step(sc, dt)` moves the plates and steps the loop."""
    lay = elastic_layout(templates, particles)
    ps = ParticleSystem(len(lay["positions"]), lay["radius"], context=context)
    ps.uploadParticles(lay["positions"], invMass=lay["inv_mass"])
    comp = softbody_compliance(softness, lay["per_body"])
    soft = SoftbodySystem(ps, [dict(b, compliance=comp) for b in lay["bodies"]], shapeMatching="local",
                          selfCollision=True)
    walls = PrimitiveSet(ps)
    tank_friction = dict(muS=0.08, muK=0.04)
    walls.addPlane((0, 1, 0), (0, 0, 0), **tank_friction)
    walls.addPlane((1, 0, 0), (-1.05, 0, 0), **tank_friction)
    walls.addPlane((-1, 0, 0), (1.05, 0, 0), **tank_friction)
    walls.addPlane((0, 0, 1), (0, 0, -0.87), **tank_friction)
    walls.addPlane((0, 0, -1), (0, 0, 0.87), **tank_friction)
    plates = [walls.addBox((0, 0, 0), (0.02, 0.21, 0.87), muS=0.1, muK=0.05) for _ in range(2)]
    n = len(lay["positions"])
    loop = SimLoop(ps, substeps=substeps or scaled_substeps(6, n), iterations=iterations, gravity=(0.0, -gravity, 0.0),
                   materials=[soft], colliders=[walls], contact={"muS": 0.12, "muK": 0.06, "maxContacts": 20 * n})
    sc = dict(ps=ps, soft=soft, walls=walls, loop=loop, layout=lay, plates=plates, time=0.0)
    _place_plates(sc, 0.0)
    return sc


def _place_plates(sc, dt):
    d = elastic_plate_distance(sc["time"])
    for side, slot in zip((-1.0, 1.0), sc["plates"]):
        # the panel's box sits 0.2 m above its group origin (elastic.ts:183-196)
        sc["walls"].setPose(slot, (side * d, 0.2, 0.0), (0.0, 0.0, 0.0, 1.0), dt)


def step_elastic_scene(sc, dt):
    """One frame of the elastic studies: plates follow the squeeze cycle, then the loop steps."""
    sc["time"] += dt
    _place_plates(sc, dt)
    sc["loop"].step(dt)


def dispose_scene(sc):
    """Release a scene built by build_mixed_scene or build_dam_break, loop first."""
    for key in ("loop", "soft", "fluid", "walls", "ps"):
        obj = sc.get(key)
        if obj is not None:
            obj.dispose()


def log_uniform_counts(count, lo, hi, seed):
    """`count` particle budgets log-uniform in [lo, hi] from `seed`."""
    rng = np.random.default_rng(seed)
    return np.exp(rng.uniform(np.log(lo), np.log(hi), count)).astype(np.int64)


def sort_points(n, distribution, seed=0x5EED):
    """Isolated-sort inputs in [0,1)^3: 'uniform' or 64 Gaussian 'blobs' (sigma 0.03)."""
    if distribution == "uniform":
        u = lcg_uniform(seed, 3 * n).reshape(n, 3)
        pts = u
    elif distribution == "blobs":
        rng = np.random.default_rng(seed)
        centers = rng.random((64, 3))
        per = n // 64
        out = []
        for b in range(64):
            pts_b = np.empty((0, 3))
            need = per if b < 63 else n - per * 63
            while len(pts_b) < need:
                cand = centers[b] + 0.03 * rng.standard_normal((2 * need, 3))
                ok = np.all((cand >= 0.0) & (cand < 1.0), axis=1)
                pts_b = np.concatenate([pts_b, cand[ok]])
            out.append(pts_b[:need])
        pts = np.concatenate(out)
    else:
        raise ValueError(distribution)
    pos = np.zeros((n, 4), np.float32)
    pos[:, :3] = pts.astype(np.float32)
    return pos
