"""Shared test helpers: building the same scene on the GPU and in the oracle."""
import numpy as np

import oracle as orc
from particle_fluids import scenes


def oracle_dam_break(positions, radius, substeps, iterations, precision="f32", H=None, effects=None):
    """The dam break of particle_fluids.scenes in the CPU oracle."""
    n = len(positions)
    spacing = 2 * radius
    h = 2 * spacing
    if H is None:
        H = 1
        while H < 2 * n:
            H <<= 1
        H = min(H, 1 << 20)
    fx = dict(scenes.FLUID_DEFAULTS) if effects is None else effects
    s = orc.Scene(n, radius, precision)
    s.set_particles(positions)
    s.set_loop(substeps, iterations, (0.0, -9.81, 0.0), h, H)
    s.add_fluid(0, n, 1000.0, spacing, h, 1e-4, viscosity=fx.get("viscosity"), vorticity=fx.get("vorticity"),
                surface_tension=fx.get("surfaceTension"), adhesion=fx.get("adhesion"))
    mn, mx = scenes.CONTAINER[:3], scenes.CONTAINER[3:]
    for nrm, which in scenes.PLANES:
        s.add_plane(nrm, mn if which == "min" else mx)
    return s


def rel_err(a, b):
    """max |a - b| / max |b| (inf-norm relative error)."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    d = np.abs(a - b).max()
    s = np.abs(b).max()
    return float(d / s) if s > 0 else float(d)


def boundary_layout(radius=0.01, plate_w=0.0):
    """A fluid block above a two-layer plate of boundary particles.

    Returns (positions (N,3) f32, inv_mass (N,) f32, n_fluid, n_plate). The
    fluid's inverse mass is set by the fluid itself; plate particles get
    plate_w (0 pins them)."""
    s = 2.0 * radius
    fluid = [(x, y, z) for x in np.arange(-0.19, 0.19 + 1e-9, s) for y in np.arange(0.09, 0.27 + 1e-9, s)
             for z in np.arange(-0.09, 0.09 + 1e-9, s)]
    plate = [(x, y, z) for x in np.arange(-0.29, 0.29 + 1e-9, s) for y in (0.045, 0.065)
             for z in np.arange(-0.15, 0.15 + 1e-9, s)]
    pos = np.array(fluid + plate, np.float32)
    w = np.zeros(len(pos), np.float32)
    w[len(fluid):] = plate_w
    return pos, w, len(fluid), len(plate)


def oracle_boundary_scene(pos, w, nf, nb, radius, dynamic, substeps=2, iterations=4, precision="f32",
                          gravity=(0.0, -9.81, 0.0), effects=None, planes=True):
    """The boundary scene in the CPU oracle (fluid [0, nf), boundary [nf, nf + nb))."""
    n = len(pos)
    spacing = 2 * radius
    h = 2 * spacing
    H = 1
    while H < 2 * n:
        H <<= 1
    fx = dict(scenes.FLUID_DEFAULTS) if effects is None else effects
    s = orc.Scene(n, radius, precision)
    s.set_particles(pos, w=w)
    s.set_loop(substeps, iterations, gravity, h, min(H, 1 << 20))
    s.add_fluid(0, nf, 1000.0, spacing, h, 1e-4, viscosity=fx.get("viscosity"), vorticity=fx.get("vorticity"),
                surface_tension=fx.get("surfaceTension"), adhesion=fx.get("adhesion"))
    s.add_boundary(0, nf, nb, dynamic)
    if planes:
        mn, mx = scenes.CONTAINER[:3], scenes.CONTAINER[3:]
        for nrm, which in scenes.PLANES:
            s.add_plane(nrm, mn if which == "min" else mx)
    return s


def voxel_grid(nx, ny, nz, radius):
    """An nx*ny*nz particle block at spacing 2r centred on the origin, with its
    6-face edge list (i < j), as the reference's local shape-matching tests
    build it (tests/analytical/softbody/softbody-implicit.gpu.test.ts:24-62)."""
    sp = 2.0 * radius
    ox, oy, oz = -((nx - 1) * sp) / 2, -((ny - 1) * sp) / 2, -((nz - 1) * sp) / 2
    idx = lambda ix, iy, iz: iz * nx * ny + iy * nx + ix
    rest = np.zeros((nx * ny * nz, 3), np.float32)
    edges = []
    for iz in range(nz):
        for iy in range(ny):
            for ix in range(nx):
                i = idx(ix, iy, iz)
                rest[i] = (ox + ix * sp, oy + iy * sp, oz + iz * sp)
                if ix + 1 < nx:
                    edges += [i, idx(ix + 1, iy, iz)]
                if iy + 1 < ny:
                    edges += [i, idx(ix, iy + 1, iz)]
                if iz + 1 < nz:
                    edges += [i, idx(ix, iy, iz + 1)]
    return rest, np.array(edges, np.uint32)


def rest_offsets(rest, weights=None):
    """Rest offsets from the mass-weighted rest centre, as SoftbodySystem
    stores them (f32 of the double difference, SoftbodySystem.ts:178-185)."""
    r = np.asarray(rest, np.float64)
    w = np.ones(len(r)) if weights is None else np.asarray(weights, np.float64)
    c = (r * w[:, None]).sum(0) / w.sum()
    return (r - c).astype(np.float32), np.asarray(w, np.float32)


def oracle_local_body(positions, velocities, inv_mass, rest, edges, compliance, radius, substeps, iterations,
                      gravity, precision="f32"):
    """One locally shape-matched body owning every particle, in the oracle."""
    n = len(positions)
    s = orc.Scene(n, radius, precision)
    s.set_particles(positions, velocities, inv_mass)
    s.set_loop(substeps, iterations, gravity, 4 * radius, 64)
    off, wts = rest_offsets(rest)
    s.add_body(0, n, compliance, off, wts)
    s.add_body_edges(edges)
    return s


def quat_to_mat(q):
    q = np.asarray(q, np.float64)
    x, y, z, w = q / np.linalg.norm(q)
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                     [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                     [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])


def oracle_mixed_scene(layout, substeps, iterations, H, precision="f32", contact=None):
    """A scene of particle_fluids.scenes.mixed_layout (dam break plus globally
    shape-matched cubes in the extended tank) in the CPU oracle. `contact` as
    scenes.build_mixed_scene: None turns contacts on when there are cubes, and
    each cube then has its own collision group (1, 2, ... in body order, as
    the SimLoop hands them out)."""
    pos, nf, r = layout["positions"], layout["fluid_count"], layout["radius"]
    n = len(pos)
    spacing = 2 * r
    h = 2 * spacing
    fx = scenes.FLUID_DEFAULTS
    s = orc.Scene(n, r, precision)
    w = layout["inv_mass"]
    s.set_particles(pos, None, w)
    s.set_loop(substeps, iterations, (0.0, -9.81, 0.0), h, H)
    s.add_fluid(0, nf, 1000.0, spacing, h, 1e-4, viscosity=fx["viscosity"], vorticity=fx["vorticity"],
                surface_tension=fx["surfaceTension"], adhesion=fx["adhesion"])
    for b in layout["bodies"]:
        st, c = b["range"]
        inv = np.asarray(w[st:st + c], np.float64)
        rest = pos[st:st + c].astype(np.float64)
        # massWeights (SoftbodySystem.ts:366-372): 1 for equal masses, else m / mean(m)
        m = np.where(inv > 0, 1.0 / np.where(inv > 0, inv, 1.0), 0.0)
        dyn = m[m > 0]
        if np.all(dyn == dyn[0]):
            wts = np.ones(c, np.float32)
        else:
            wts = np.where(m > 0, m / dyn.mean(), 1.0).astype(np.float32)
        # weighted rest centre summed in order, in double (SoftbodySystem.ts:374-386)
        ctr = np.zeros(3)
        for p, wk in zip(rest, wts.astype(np.float64)):
            ctr += p * wk
        ctr /= float(wts.astype(np.float64).sum())
        s.add_body(st, c, b["compliance"], (rest - ctr).astype(np.float32), wts)
    mn, mx = layout["container"][:3], layout["container"][3:]
    for nrm, which in scenes.PLANES:
        s.add_plane(nrm, mn if which == "min" else mx)
    if layout["bodies"]:
        groups = np.zeros(n, np.uint32)
        for b, body in enumerate(layout["bodies"]):
            st, c = body["range"]
            groups[st:st + c] = b + 1
        s.set_groups(groups)
    if contact is None and layout["bodies"]:
        soft_count = sum(b["range"][1] for b in layout["bodies"])
        contact = dict(scenes.SIMULATION_CONTACT, maxContacts=8 * soft_count)
    elif contact is None:
        contact = False
    if contact:
        c = {} if contact is True else contact
        s.set_contact(c.get("muS", 0.5), c.get("muK", 0.4), c.get("maxContacts"))
    return s


# ---------------------------------------------------------------------------
# Posed primitives: PrimitiveSet.attach / update / place (PrimitiveSet.ts:
# 211-234, 259-285, 469-486) restated in double precision, giving the values
# the oracle's set_prim takes. Quaternions are (x, y, z, w) as in three.js.
# ---------------------------------------------------------------------------
def qmul(a, b):
    ax, ay, az, aw = a
    bx, by, bz, bw = b
    return np.array([ax * bw + aw * bx + ay * bz - az * by, ay * bw + aw * by + az * bx - ax * bz,
                     az * bw + aw * bz + ax * by - ay * bx, aw * bw - ax * bx - ay * by - az * bz])


def qconj(q):
    return np.array([-q[0], -q[1], -q[2], q[3]])


def qapply(q, v):
    # Vector3.applyQuaternion: t = 2 q.xyz x v; v + w t + q.xyz x t
    u = np.asarray(q[:3], np.float64)
    t = 2.0 * np.cross(u, v)
    return np.asarray(v, np.float64) + q[3] * t + np.cross(u, t)


class PosedPrimitive:
    """One attached box or capsule: data0/data1/velocity/rotation/spin as
    PrimitiveSet keeps them, driven by world poses per step."""
    KINDS = {"box": 2, "capsule": 3}

    def __init__(self, kind, data0, data1, mus=0.5, muk=0.4):
        self.kind = self.KINDS[kind]
        self.d0 = np.array(data0, np.float64)
        self.d1 = np.array(data1, np.float64)
        self.d1[3] = mus
        self.vel = np.array([0.0, 0.0, 0.0, muk])
        self.rot = np.array([0.0, 0.0, 0.0, 1.0])
        self.spin = np.zeros(3)
        self.flags = 0
        self.attached = False

    def pose(self, position, rotation, dt=0.0):
        p = np.asarray(position, np.float64)
        q = np.asarray(rotation, np.float64)
        if not self.attached:
            if self.kind == 3:
                self.direction = (self.d1[:3] - self.d0[:3]) * 0.5
            else:
                self.direction = np.zeros(3)
            self.offset = np.array([0.0, 0.0, 0.0, 1.0]) if self.kind == 2 else qconj(q)
            self.flags |= 2
            self.attached = True
        elif dt > 0:
            self.vel[:3] = (p - self.prev) / dt
            turn = qmul(q, qconj(self.prev_rot))
            if turn[3] < 0:
                turn = -turn
            s = np.hypot(turn[0], np.hypot(turn[1], turn[2]))
            rate = (2 * np.arctan2(s, turn[3])) / (s * dt) if s > 1e-12 else 0.0
            self.spin = turn[:3] * rate
        self.rot = qmul(q, self.offset)
        d = qapply(self.rot, self.direction)
        if self.kind == 3:
            self.d0[:3] = p - d
            self.d1[:3] = p + d
        else:
            self.d0[:3] = p
        self.prev = p.copy()
        self.prev_rot = q.copy()

    def oracle_args(self):
        return self.kind, self.flags, self.d0, self.d1, self.vel, self.rot, self.spin


CONTACT_RADIUS_EXPANSION = 1.1  # candidate radius 2r * 1.1 (generate.ts:19)


def hash_table_size(n):
    """H = min(nextPow2(2n), 2^20) (HashGrid.ts:122-124)."""
    H = 1
    while H < 2 * n:
        H <<= 1
    return min(H, 1 << 20)


def oracle_contact_scene(pos, radius, substeps, iterations, gravity=(0.0, -9.81, 0.0), vel=None, w=None,
                         groups=None, mus=0.5, muk=0.4, max_contacts=None, quiet=(), precision="f32", cell=None):
    """Particles with SimLoop contacts on (loop.ts:346-418) and no materials
    in the oracle; the grid's cell is the contact candidate radius."""
    n = len(pos)
    s = orc.Scene(n, radius, precision)
    s.set_particles(pos, vel, w)
    if cell is None:
        cell = 2 * radius * CONTACT_RADIUS_EXPANSION
    s.set_loop(substeps, iterations, gravity, cell, hash_table_size(n))
    if groups is not None:
        s.set_groups(groups)
    s.set_contact(mus, muk, max_contacts, quiet)
    return s


def brute_force_pairs(pos, w, groups, radius, emitter):
    """Every pair the contact search must find (generate.ts:39-88), from all
    pairs rather than the grid walk: at least one end emits, not one non-zero
    group, not both pinned, closer than 2r * 1.1 (distance in f32)."""
    p = np.asarray(pos, np.float32)
    cr = 2.0 * radius * CONTACT_RADIUS_EXPANSION
    cr2 = np.float32(cr * cr)
    out = set()
    n = len(p)
    for i in range(n):
        for j in range(i + 1, n):
            if not (emitter[i] or emitter[j]):
                continue
            if groups[i] != 0 and groups[i] == groups[j]:
                continue
            if w[i] <= 0 and w[j] <= 0:
                continue
            d = p[i] - p[j]
            d2 = np.float32(np.float32(d[0] * d[0]) + np.float32(d[1] * d[1])) + np.float32(d[2] * d[2])
            if d2 < cr2:
                out.add((i, j))
    return out


def floating_cubes_layout(budget=3000, side=5, seed=11):
    """A dam-break water block with two soft cubes resting on its surface,
    each overlapping the top water layer and the other cube by a tenth of a
    spacing: contacts (and their stabilisation) start on the first substep."""
    pos, r = scenes.dam_break_layout(budget)
    pos = scenes.jitter(pos, 2 * r, seed)
    nf = len(pos)
    sp = 2 * r
    lo, hi = pos.min(0).astype(np.float64), pos.max(0).astype(np.float64)
    half = (side - 1) / 2 * sp
    xc = 0.5 * (lo[0] + hi[0])
    yc = hi[1] + 0.9 * sp + half
    za = lo[2] + half + 2 * sp
    cubes, bodies = [], []
    for k, zc in enumerate((za, za + (side - 1) * sp + 0.9 * sp)):
        cp, surf = scenes.particle_cube((xc, yc, zc), side, sp)
        bodies.append(dict(range=(nf + k * side ** 3, side ** 3), surfaceCount=surf,
                           compliance=scenes.softbody_compliance(0.5, side ** 3)))
        cubes.append(cp)
    allpos = np.concatenate([pos.astype(np.float64)] + cubes).astype(np.float32)
    w = np.empty(len(allpos), np.float32)
    w[:nf] = 0.0
    w[nf:] = 1.0 / (scenes.CUBE_DENSITY * sp ** 3)
    return dict(positions=allpos, inv_mass=w, fluid_count=nf, bodies=bodies, radius=r,
                container=tuple(scenes.CONTAINER))


def oracle_elastic_scene(sc, precision="f32"):
    """The GPU elastic studies scene `sc` (scenes.build_elastic_scene) in the
    oracle: the same particles, locally shape-matched bodies with their
    edges, contacts with every collision group 0 (selfCollision), the tank's
    planes and the two plates as posed boxes. Returns (scene, plates)."""
    lay = sc["layout"]
    pos, w = lay["positions"], lay["inv_mass"]
    n = len(pos)
    s = orc.Scene(n, lay["radius"], precision)
    s.set_particles(pos, None, w)
    info = sc["loop"].info()
    s.set_loop(info["substeps"], info["iterations"], (0.0, -6.0, 0.0), info["cellSize"], info["hashTableSize"])
    comp = scenes.softbody_compliance(0.9, lay["per_body"])
    for b in lay["bodies"]:
        st, c = b["range"]
        off, wts = rest_offsets(pos[st:st + c])
        s.add_body(st, c, comp, off, wts)
        s.add_body_edges(np.asarray(b["edges"], np.uint32))
    s.set_contact(0.12, 0.06, 20 * n)
    tank = dict(mus=0.08, muk=0.04)
    s.add_plane((0, 1, 0), (0, 0, 0), **tank)
    s.add_plane((1, 0, 0), (-1.05, 0, 0), **tank)
    s.add_plane((-1, 0, 0), (1.05, 0, 0), **tank)
    s.add_plane((0, 0, 1), (0, 0, -0.87), **tank)
    s.add_plane((0, 0, -1), (0, 0, 0.87), **tank)
    plates = [PosedPrimitive("box", [0, 0, 0, 0], [0.02, 0.21, 0.87, 0], mus=0.1, muk=0.05) for _ in range(2)]
    for k, pl in enumerate(plates):
        pl.pose((-1.0 if k == 0 else 1.0) * scenes.elastic_plate_distance(0.0) * np.array([1, 0, 0])
                + np.array([0, 0.2, 0]), (0, 0, 0, 1))
        s.add_prim(*pl.oracle_args())
    return s, plates


def step_oracle_elastic(s, plates, time, dt):
    """One frame of the elastic studies in the oracle (plates first, as
    scenes.step_elastic_scene does); `time` is the frame's end time."""
    d = scenes.elastic_plate_distance(time)
    for k, pl in enumerate(plates):
        pl.pose(((-1.0 if k == 0 else 1.0) * d, 0.2, 0.0), (0, 0, 0, 1), dt)
        s.set_prim(5 + k, *pl.oracle_args())
    s.step(dt)
