"""Simulation facade (src/simulation/Simulation.ts:301-1144) over the native solver.

Scene construction follows the reference: particles are sized from a budget
(fitBudget, Simulation.ts:957-1002) or a radius, fluid boxes are filled at
spacing 2r inside the container shrunk by r (layout, 1044-1104; fillBox,
layout.ts:7-14), points inside obstacles are dropped (clearOf, layout.ts:83-124),
and the container becomes five planes with an open top (777-787). Rendering,
meshes, cloth and smoke are outside the native solver.
"""
import math

import numpy as np

from . import (FluidSystem, ParticleSystem, PFError, PrimitiveSet, SimLoop, ViscositySolver, default_context, fill_box,
               radius_for_budget)


def _box_intersect(a, b):
    lo = [max(a[i], b[i]) for i in range(3)]
    hi = [min(a[3 + i], b[3 + i]) for i in range(3)]
    return lo + hi


def _box_volume(b):
    s = [b[3 + i] - b[i] for i in range(3)]
    if any(v < 0 for v in s):
        return 0.0
    return s[0] * s[1] * s[2]


def _quat_apply(q, v):
    x, y, z, w = q
    vx, vy, vz = v[..., 0], v[..., 1], v[..., 2]
    ix = w * vx + y * vz - z * vy
    iy = w * vy + z * vx - x * vz
    iz = w * vz + x * vy - y * vx
    iw = -x * vx - y * vy - z * vz
    return np.stack([ix * w + iw * -x + iy * -z - iz * -y, iy * w + iw * -y + iz * -x - ix * -z,
                     iz * w + iw * -z + ix * -y - iy * -x], axis=-1)


def _euler_xyz_quat(e):
    # three.js Quaternion.setFromEuler, order 'XYZ'
    x, y, z = e
    c1, c2, c3 = math.cos(x / 2), math.cos(y / 2), math.cos(z / 2)
    s1, s2, s3 = math.sin(x / 2), math.sin(y / 2), math.sin(z / 2)
    return (s1 * c2 * c3 + c1 * s2 * s3, c1 * s2 * c3 - s1 * c2 * s3, c1 * c2 * s3 + s1 * s2 * c3, c1 * c2 * c3 - s1 * s2 * s3)


class Simulation:
    """new Simulation({particles | particleRadius, container, closed, gravity, substeps}).

    ``iterations`` is an addition: the reference fixes 2 (Simulation.ts:830-837);
    the default here is 2 as well.
    """

    def __init__(self, particles=None, particleRadius=None, container=None, closed=False, gravity=None,
                 substeps=None, iterations=2, context=None):
        if (particles is None) == (particleRadius is None):
            raise PFError(1, "Simulation: give exactly one of `particles` or `particleRadius`")
        if particles is not None and not (isinstance(particles, int) and particles > 0):
            raise PFError(1, f"Simulation: particles must be a positive integer, got {particles}")
        if particleRadius is not None and not (particleRadius > 0 and math.isfinite(particleRadius)):
            raise PFError(1, f"Simulation: particleRadius must be a positive number of metres, got {particleRadius}")
        self.options = dict(particles=particles, particleRadius=particleRadius, container=container, closed=closed,
                            substeps=substeps, iterations=iterations)
        self.context = context or default_context()
        self.gravity = list(gravity) if gravity is not None else [0.0, -9.81, 0.0]
        self.radius = particleRadius or 0.0
        self.fluids = []
        self.colliders = []
        self.started = False
        self.system = None
        self.simLoop = None
        self.walls = None
        self.fluidSystems = []
        self.fluidRanges = []

    # -- scene description (before the first step) ------------------------
    def _not_started(self, what):
        if self.started:
            raise PFError(1, f"Simulation: {what} must happen before the first step() or start()")

    def addFluid(self, box, viscosity=0.01, surfaceTension=0.1, vorticity=0.02, adhesion=0.1, thickness=None):
        """addFluid({box, viscosity, surfaceTension, vorticity, adhesion, thickness}) (Simulation.ts:342-368).
        `thickness` adds a ViscositySolver (honey is about 20) with round(16 cbrt(N / 1e4))
        sweeps per substep (Simulation.ts:740-745)."""
        self._not_started("addFluid")
        if box is None:
            raise PFError(1, "addFluid: give either `box` or `mesh`")
        entry = dict(box=[float(v) for v in box], viscosity=viscosity, surfaceTension=surfaceTension,
                     vorticity=vorticity, adhesion=adhesion, thickness=thickness)
        self.fluids.append(entry)
        return entry

    def addFloor(self, height=0.0, friction=0.5):
        self._not_started("adding obstacles")
        self.colliders.append(dict(kind="floor", height=float(height), friction=float(friction)))

    def addSphere(self, radius, center=(0.0, 0.0, 0.0), friction=0.5):
        self._not_started("adding obstacles")
        self.colliders.append(dict(kind="sphere", center=[float(v) for v in center], radius=float(radius), friction=float(friction)))

    def addBox(self, size, center=(0.0, 0.0, 0.0), rotation=(0.0, 0.0, 0.0), friction=0.5):
        self._not_started("adding obstacles")
        self.colliders.append(dict(kind="box", center=[float(v) for v in center], half=[0.5 * float(v) for v in size],
                                   rotation=_euler_xyz_quat(rotation), friction=float(friction)))

    def addCapsule(self, start, end, radius, friction=0.5):
        self._not_started("adding obstacles")
        self.colliders.append(dict(kind="capsule", a=[float(v) for v in start], b=[float(v) for v in end],
                                   radius=float(radius), friction=float(friction)))

    # -- layout ------------------------------------------------------------
    def _obstacles(self):
        obs = []
        for d in self.colliders:
            if d["kind"] == "sphere":
                c, r = np.array(d["center"]), d["radius"]
                obs.append(lambda p, c=c, r=r: np.linalg.norm(p - c, axis=1) - r)
            elif d["kind"] == "box":
                c, half, q = np.array(d["center"]), np.array(d["half"]), d["rotation"]
                qi = (-q[0], -q[1], -q[2], q[3])

                def box(p, c=c, half=half, qi=qi):
                    loc = _quat_apply(qi, p - c)
                    qd = np.abs(loc) - half
                    outside = np.linalg.norm(np.maximum(qd, 0), axis=1)
                    return outside + np.minimum(qd.max(axis=1), 0)
                obs.append(box)
            elif d["kind"] == "capsule":
                a, b, r = np.array(d["a"]), np.array(d["b"]), d["radius"]
                ab = b - a
                ll = max(float(ab @ ab), 1e-12)

                def cap(p, a=a, ab=ab, ll=ll, r=r):
                    t = np.clip(((p - a) @ ab) / ll, 0, 1)
                    return np.linalg.norm(p - (a + ab * t[:, None]), axis=1) - r
                obs.append(cap)
            elif d["kind"] == "floor":
                obs.append(lambda p, h=d["height"]: p[:, 1] - h)
        return obs

    def _layout(self, r, obstacles):
        container = self.options["container"]
        spacing = 2 * r
        init = []
        ranges = []
        for f in self.fluids:
            box = f["box"]
            if container is not None:
                shrunk = [container[i] + r for i in range(3)] + [container[3 + i] - r for i in range(3)]
                box = _box_intersect(box, shrunk)
            pts = fill_box(box[:3], box[3:], spacing)
            if obstacles and len(pts):
                p64 = pts.astype(np.float64)
                keep = np.ones(len(pts), bool)
                for ob in obstacles:
                    keep &= ~(ob(p64) < spacing)
                pts = pts[keep]
            start = sum(len(x) for x in init)
            init.append(pts)
            ranges.append((start, len(pts)))
        pos = np.concatenate(init) if init else np.zeros((0, 3), np.float32)
        return pos, ranges

    def _fit_budget(self, budget, obstacles):
        container = self.options["container"]
        volume = 0.0
        for f in self.fluids:
            b = _box_intersect(f["box"], container) if container is not None else f["box"]
            volume += _box_volume(b)
        guess = radius_for_budget(volume, 0.0, 2.2, budget)
        r, best = guess, None
        for _ in range(12):
            pos, ranges = self._layout(r, obstacles)
            count = len(pos)
            if count <= budget and (best is None or count > len(best[1])):
                best = (r, pos, ranges)
            if count <= budget and count >= 0.97 * budget:
                break
            r *= float(np.cbrt(max(count, 1) / (0.985 * budget)))
        for _ in range(100):
            if best is not None:
                break
            r *= 1.05
            pos, ranges = self._layout(r, obstacles)
            if len(pos) <= budget:
                best = (r, pos, ranges)
        if best is None:
            raise PFError(1, f"Simulation: the scene can't fit in {budget} particles")
        return best

    # -- lifecycle ---------------------------------------------------------
    def start(self):
        if self.started:
            return
        if not self.fluids:
            raise PFError(1, "Simulation: add a fluid, smoke, soft body, or cloth before stepping")
        obstacles = self._obstacles()
        if self.options["particleRadius"] is not None:
            r = self.options["particleRadius"]
            pos, ranges = self._layout(r, obstacles)
        else:
            r, pos, ranges = self._fit_budget(self.options["particles"], obstacles)
        for _, c in ranges:
            if c == 0:
                raise PFError(1, "addFluid: the fluid has no room; check its box and the container")
        self.radius = r
        ps = ParticleSystem(len(pos), r, context=self.context)
        ps.uploadParticles(pos)
        materials = []
        fluid_systems = []
        for f, rng in zip(self.fluids, ranges):
            fl = FluidSystem(ps, range=rng, viscosity=f["viscosity"], surfaceTension=f["surfaceTension"],
                             vorticity=f["vorticity"], adhesion=f["adhesion"])
            materials.append(fl)
            fluid_systems.append(fl)
            if f.get("thickness") is not None:
                # Math.round(16 * Math.cbrt(count / 10000)): halves round up, as in JS
                sweeps = int(math.floor(16.0 * float(np.cbrt(rng[1] / 10000.0)) + 0.5))
                materials.append(ViscositySolver(fl, f["thickness"], iterations=sweeps))
        walls = PrimitiveSet(ps)
        container = self.options["container"]
        if container is not None:
            mn, mx = container[:3], container[3:]
            walls.addPlane((0, 1, 0), mn)
            walls.addPlane((1, 0, 0), mn)
            walls.addPlane((-1, 0, 0), mx)
            walls.addPlane((0, 0, 1), mn)
            walls.addPlane((0, 0, -1), mx)
            if self.options["closed"]:
                walls.addPlane((0, -1, 0), mx)
        for d in self.colliders:
            muS, muK = d["friction"], d["friction"] * 0.8
            if d["kind"] == "floor":
                walls.addPlane((0, 1, 0), (0, d["height"], 0), muS, muK)
            elif d["kind"] == "sphere":
                walls.addSphere(d["center"], d["radius"], muS=muS, muK=muK)
            elif d["kind"] == "box":
                walls.addBox(d["center"], d["half"], rotation=d["rotation"], muS=muS, muK=muK)
            elif d["kind"] == "capsule":
                walls.addCapsule(d["a"], d["b"], d["radius"], muS=muS, muK=muK)
        substeps = self.options["substeps"]
        if substeps is None:
            substeps = min(24, math.ceil(4 * max(1.0, 0.018 / r)))
        colliders = [walls] if walls.count > 0 else []
        self.simLoop = SimLoop(ps, substeps=substeps, iterations=self.options["iterations"], gravity=self.gravity,
                               materials=materials, colliders=colliders)
        self.system, self.walls, self.fluidSystems, self.fluidRanges = ps, walls, fluid_systems, ranges
        self.materials = materials
        self.started = True

    def step(self, dt=1.0 / 60.0):
        self.start()
        self.simLoop.setGravity(self.gravity)
        self.simLoop.step(dt)

    @property
    def particleRadius(self):
        return self.radius

    @property
    def particleCount(self):
        return self.system.capacity if self.system else 0

    @property
    def particleSystem(self):
        if not self.system:
            raise PFError(1, "Simulation.particleSystem is created on the first step")
        return self.system

    @property
    def loop(self):
        if not self.simLoop:
            raise PFError(1, "Simulation.loop is created on the first step")
        return self.simLoop

    def dispose(self):
        if self.simLoop:
            self.simLoop.dispose()
        for m in getattr(self, "materials", []):
            if isinstance(m, ViscositySolver):
                m.dispose()
        for f in self.fluidSystems:
            f.dispose()
        if self.walls:
            self.walls.dispose()
        if self.system:
            self.system.dispose()
        self.simLoop = self.system = self.walls = None
        self.fluidSystems = []
