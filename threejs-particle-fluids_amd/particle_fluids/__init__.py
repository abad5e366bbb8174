"""Python host mirror of threejs-particle-fluids' solver API over libpf.so.

Class, method and option names follow the reference (src/core/*,
src/fluids/FluidSystem.ts, src/softbody/SoftbodySystem.ts,
src/simulation/Simulation.ts) so code written against it reads the same;
errors carry the reference's messages. Every call goes to the HIP library:
there is no CPU fallback.
"""
import ctypes as C
import math
from collections import namedtuple

import numpy as np

from . import _lib
from ._lib import PFError, call, load

__all__ = [
    "PFError", "Context", "ParticleSystem", "HashGrid", "FluidSystem", "PrimitiveSet",
    "SoftbodySystem", "SDFCollider", "SimLoop", "Batch", "Simulation", "fill_box", "fit_budget",
    "radius_for_budget", "NoSelfContacts", "ContactBuffer", "SurfaceField", "ViscositySolver",
]

P = C.POINTER
f32p = P(C.c_float)
u32p = P(C.c_uint32)
f64p = P(C.c_double)


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _ptr(a, t):
    return None if a is None else a.ctypes.data_as(t)


def _d3(v):
    return (C.c_double * 3)(*[float(x) for x in v])


def _d4(v):
    return (C.c_double * 4)(*[float(x) for x in v])


def _range(r, default_count):
    if r is None:
        return 0, default_count
    if isinstance(r, dict):
        return int(r["start"]), int(r["count"])
    return int(r[0]), int(r[1])


def upload_arrays(positions, velocities=None, invMass=None, collisionGroup=None):
    """uploadParticles' arrays as the C ABI reads them: n xyz positions, n xyz
    velocities, n inverse masses and n groups (one value broadcasts). Arrays
    that do not cover the n particles raise ValueError, since the library
    reads n entries of each."""
    pos = _f32(positions).reshape(-1, 3)
    n = pos.shape[0]
    vel = None if velocities is None else _f32(velocities).reshape(-1, 3)
    if vel is not None and vel.shape[0] != n:
        raise ValueError(f"ParticleSystem.uploadParticles: {vel.shape[0]} velocities for {n} particles")
    try:
        im = None if invMass is None else _f32(np.broadcast_to(invMass, (n,)))
        grp = None if collisionGroup is None else np.ascontiguousarray(np.broadcast_to(collisionGroup, (n,)),
                                                                        dtype=np.uint32)
    except ValueError:
        raise ValueError(f"ParticleSystem.uploadParticles: invMass and collisionGroup need one value or {n} values") \
            from None
    return pos, vel, im, grp


class Context:
    """One GPU and one HIP stream (replaces the WebGPU renderer handle)."""

    def __init__(self, device=0):
        h = C.c_void_p()
        call("pf_context_create", int(device), C.byref(h))
        self.handle = h

    def synchronize(self):
        call("pf_context_synchronize", self.handle)

    def timer_begin(self):
        call("pf_context_timer_begin", self.handle)

    def timer_end(self):
        ms = C.c_float()
        call("pf_context_timer_end", self.handle, C.byref(ms))
        return ms.value

    def set_profiling(self, on):
        call("pf_context_set_profiling", self.handle, 1 if on else 0)

    def profile(self):
        """{kernel name: (launches, total ms)} since the last reset."""
        n = C.c_uint32()
        call("pf_context_profile_count", self.handle, C.byref(n))
        out = {}
        buf = C.create_string_buffer(128)
        for i in range(n.value):
            launches = C.c_uint32()
            total = C.c_double()
            call("pf_context_profile_entry", self.handle, i, buf, 128, C.byref(launches), C.byref(total))
            out[buf.value.decode()] = (launches.value, total.value)
        return out

    def profile_reset(self):
        call("pf_context_profile_reset", self.handle)

    def dispose(self):
        if self.handle:
            call("pf_context_destroy", self.handle)
            self.handle = None


_default = None


def default_context():
    global _default
    if _default is None:
        _default = Context(0)
    return _default


ParticleSnapshot = namedtuple(
    "ParticleSnapshot",
    "capacity positions predictedPositions velocities invMass rotation predictedRotation angularVelocity",
)
HashGridSnapshot = namedtuple(
    "HashGridSnapshot",
    "capacity hashTableSize hashTableSizePadded cellIndex counts cellStart cellEnd sortedIndices",
)


class ParticleSystem:
    """src/core/particles.ts:52-258. Positions are (N, 3) arrays."""

    def __init__(self, capacity, particleRadius, context=None, evictable=False):
        self.context = context or default_context()
        h = C.c_void_p()
        fn = "pf_particles_create_evictable" if evictable else "pf_particles_create"
        call(fn, self.context.handle, int(capacity), float(particleRadius), C.byref(h))
        self.handle = h
        self.evictable = bool(evictable)
        self.capacity = int(capacity)
        self.particleRadius = float(particleRadius)

    def uploadParticles(self, positions, velocities=None, invMass=None, collisionGroup=None, start=0):
        pos, vel, im, grp = upload_arrays(positions, velocities, invMass, collisionGroup)
        call("pf_particles_upload", self.handle, int(start), pos.shape[0], _ptr(pos, f32p), _ptr(vel, f32p),
             _ptr(im, f32p), _ptr(grp, u32p))

    def setInvMass(self, range, invMass):
        s, c = _range(range, self.capacity)
        call("pf_particles_set_inv_mass", self.handle, s, c, float(invMass))

    def setCollisionGroup(self, range, group):
        s, c = _range(range, self.capacity)
        call("pf_particles_set_collision_group", self.handle, s, c, int(group))

    def readback(self):
        n = self.capacity
        arrs = [np.zeros((n, 4), np.float32) for _ in range(3)] + [np.zeros(n, np.float32)] + [np.zeros((n, 4), np.float32) for _ in range(3)]
        snap = _lib.Snapshot(*[_ptr(a, f32p) for a in arrs])
        call("pf_particles_readback", self.handle, C.byref(snap))
        return ParticleSnapshot(n, *arrs)

    def deviceBytes(self):
        """Device bytes of an evictable scene's arena (0 otherwise)."""
        b = C.c_uint64()
        call("pf_particles_device_bytes", self.handle, C.byref(b))
        return b.value

    def readbackBoundaryVolume(self):
        """The boundaryVolume buffer (Akinci volumes of boundary particles, else 0)."""
        v = np.zeros(self.capacity, np.float32)
        call("pf_particles_readback_boundary_volume", self.handle, _ptr(v, f32p))
        return v

    def dispose(self):
        if self.handle:
            call("pf_particles_destroy", self.handle)
            self.handle = None

    @property
    def disposed(self):
        return self.handle is None


class HashGrid:
    """src/core/hashGrid/HashGrid.ts:80-286.

    ``HashGrid(particles, cellSize, ...)`` bins a ParticleSystem's predicted
    positions; ``HashGrid.standalone(capacity, cellSize, ...)`` sorts points
    uploaded with :meth:`uploadPositions`.
    """

    def __init__(self, particles, cellSize, hashTableSize=None, hashOrigin=None, _handle=None, _owner=None):
        self.particles = particles
        self._owned = _owner is not None
        self._owner = _owner
        if _handle is not None:
            self.handle = _handle
        else:
            h = C.c_void_p()
            origin = None if hashOrigin is None else _d3(hashOrigin)
            call("pf_grid_create", particles.handle, float(cellSize), int(hashTableSize or 0), origin, C.byref(h))
            self.handle = h
        self._info()

    @classmethod
    def standalone(cls, capacity, cellSize, hashTableSize=None, hashOrigin=None, context=None):
        ctx = context or default_context()
        h = C.c_void_p()
        origin = None if hashOrigin is None else _d3(hashOrigin)
        call("pf_grid_create_standalone", ctx.handle, int(capacity), float(cellSize), int(hashTableSize or 0), origin, C.byref(h))
        g = cls.__new__(cls)
        g.particles = None
        g._owned = False
        g._owner = None
        g.handle = h
        g._info()
        return g

    def _info(self):
        cap, H, Hp = C.c_uint32(), C.c_uint32(), C.c_uint32()
        call("pf_grid_info", self.handle, C.byref(cap), C.byref(H), C.byref(Hp))
        self.capacity, self.hashTableSize, self.hashTableSizePadded = cap.value, H.value, Hp.value

    def uploadPositions(self, positions):
        p = _f32(positions)
        p = p.reshape(-1, p.shape[-1] if p.ndim == 2 else 4)
        if p.shape[1] == 3:
            p = np.concatenate([p, np.zeros((p.shape[0], 1), np.float32)], axis=1)
        p = _f32(p)
        call("pf_grid_upload_positions", self.handle, _ptr(p, f32p), p.shape[0])

    def setSortedPositions(self, enabled):
        call("pf_grid_set_sorted_positions", self.handle, 1 if enabled else 0)

    def rebuild(self):
        call("pf_grid_rebuild", self.handle)

    def readback(self):
        n, Hp = self.capacity, self.hashTableSizePadded
        ci = np.zeros(n, np.uint32)
        cnt = np.zeros(Hp, np.uint32)
        cs = np.zeros(Hp, np.uint32)
        ce = np.zeros(Hp, np.uint32)
        si = np.zeros(n, np.uint32)
        call("pf_grid_readback", self.handle, _ptr(ci, u32p), _ptr(cnt, u32p), _ptr(cs, u32p), _ptr(ce, u32p), _ptr(si, u32p))
        return HashGridSnapshot(n, self.hashTableSize, Hp, ci, cnt, cs, ce, si)

    def readbackSortedKeys(self):
        k = np.zeros(self.capacity, np.uint32)
        call("pf_grid_readback_sorted_keys", self.handle, _ptr(k, u32p))
        return k

    def readbackSortedPositions(self):
        p = np.zeros((self.capacity, 4), np.float32)
        call("pf_grid_readback_sorted_positions", self.handle, _ptr(p, f32p))
        return p

    def readbackOverflow(self):
        o = C.c_int()
        call("pf_grid_readback_overflow", self.handle, C.byref(o))
        return bool(o.value)

    def queryCounts(self, points, radius):
        q = _f32(points).reshape(-1, 4)
        cand = np.zeros(q.shape[0], np.uint32)
        within = np.zeros(q.shape[0], np.uint32)
        call("pf_grid_query_counts", self.handle, _ptr(q, f32p), q.shape[0], float(radius), _ptr(cand, u32p), _ptr(within, u32p))
        return cand, within

    @property
    def hashOrigin(self):
        raise AttributeError("write-only here; use setHashOrigin")

    def setHashOrigin(self, origin):
        call("pf_grid_set_origin", self.handle, _d3(origin))

    def dispose(self):
        if self.handle and not self._owned:
            call("pf_grid_destroy", self.handle)
        self.handle = None


_FX = {"viscosity": 1, "vorticity": 2, "surfaceTension": 4, "adhesion": 8}


class FluidSystem:
    """src/fluids/FluidSystem.ts:71-322."""

    def __init__(self, particles, range=None, restDensity=None, particleSpacing=None, smoothingRadius=None,
                 compliance=None, viscosity=None, vorticity=None, surfaceTension=None, adhesion=None):
        self.particles = particles
        s, c = _range(range, 0)
        o = _lib.FluidOptions()
        o.range_start, o.range_count = s, c
        o.rest_density = float(restDensity) if restDensity is not None else 0.0
        o.particle_spacing = float(particleSpacing) if particleSpacing is not None else 0.0
        o.smoothing_radius = float(smoothingRadius) if smoothingRadius is not None else 0.0
        o.compliance = float(compliance) if compliance is not None else -1.0
        present = 0
        for name, val in (("viscosity", viscosity), ("vorticity", vorticity), ("surfaceTension", surfaceTension), ("adhesion", adhesion)):
            if val is not None:
                present |= _FX[name]
        o.present = present
        o.viscosity = float(viscosity or 0.0)
        o.vorticity = float(vorticity or 0.0)
        o.surface_tension = float(surfaceTension or 0.0)
        o.adhesion = float(adhesion or 0.0)
        h = C.c_void_p()
        call("pf_fluid_create", particles.handle, C.byref(o), C.byref(h))
        self.handle = h
        m, hh, sp, rho = C.c_double(), C.c_double(), C.c_double(), C.c_double()
        call("pf_fluid_info", h, C.byref(m), C.byref(hh), C.byref(sp), C.byref(rho))
        self.mass, self.smoothingRadius, self.particleSpacing, self.restDensity = m.value, hh.value, sp.value, rho.value
        self.neighborRadius = self.smoothingRadius
        self.range = {"start": s, "count": c or particles.capacity}

    def _get(self, key):
        v = C.c_double()
        call("pf_fluid_get", self.handle, _FX[key], C.byref(v))
        return v.value

    def _set(self, key, value):
        call("pf_fluid_set", self.handle, _FX[key], float(value))

    viscosity = property(lambda self: self._get("viscosity"), lambda self, v: self._set("viscosity", v))
    vorticity = property(lambda self: self._get("vorticity"), lambda self, v: self._set("vorticity", v))
    surfaceTension = property(lambda self: self._get("surfaceTension"), lambda self, v: self._set("surfaceTension", v))
    adhesion = property(lambda self: self._get("adhesion"), lambda self, v: self._set("adhesion", v))

    def addBoundary(self, range, dynamic=True):
        s, c = _range(range, 0)
        call("pf_fluid_add_boundary", self.handle, s, c, 1 if dynamic else 0)

    def readbackDensity(self):
        d = np.zeros(self.particles.capacity, np.float32)
        call("pf_fluid_readback_density", self.handle, _ptr(d, f32p))
        return d

    def readbackOverflow(self):
        o = C.c_int()
        call("pf_fluid_readback_overflow", self.handle, C.byref(o))
        return bool(o.value)

    def setTileUnionLimit(self, maxParticles):
        """LDS staging limit per 256-slot tile (no reference counterpart;
        results do not depend on it; 0 gathers every tile from global memory)."""
        call("pf_fluid_set_tile_union_limit", self.handle, int(maxParticles))

    def tileUnionStats(self):
        """After the last neighbour build: tiles, tiles that gathered from
        global memory, the largest and the mean staged union (no reference
        counterpart; a diagnostic of setTileUnionLimit)."""
        t, g, big = C.c_uint32(), C.c_uint32(), C.c_uint32()
        mean = C.c_double()
        call("pf_fluid_tile_union_stats", self.handle, C.byref(t), C.byref(g), C.byref(big), C.byref(mean))
        return {"tiles": t.value, "global_tiles": g.value, "largest_union": big.value, "mean_union": mean.value}

    def material(self):
        return _lib.Material(1, self.handle)

    def dispose(self):
        if self.handle:
            call("pf_fluid_destroy", self.handle)
            self.handle = None


class ViscositySolver:
    """src/fluids/ViscositySolver.ts:38-131: implicit viscosity for thick
    liquids, Jacobi sweeps of (I - dt nu L) v = v0 each substep. List it
    directly after its FluidSystem in the SimLoop's materials."""

    def __init__(self, fluid, viscosity, iterations=12):
        h = C.c_void_p()
        call("pf_viscosity_create", fluid.handle, float(viscosity), int(iterations), C.byref(h))
        self.handle = h
        self.fluid = fluid
        self.iterations = int(iterations)
        self._viscosity = float(viscosity)

    @property
    def viscosity(self):
        return self._viscosity

    @viscosity.setter
    def viscosity(self, value):
        call("pf_viscosity_set", self.handle, float(value))
        self._viscosity = float(value)

    def material(self):
        return _lib.Material(3, self.handle)

    def dispose(self):
        if self.handle:
            call("pf_viscosity_destroy", self.handle)
            self.handle = None


class SurfaceField:
    """The liquid's surface as a voxel grid (src/fluids/render/SurfaceField.ts:
    101-369), the field FluidSurfaceRenderer sphere-traces, computed headless:
    `update()` splats the committed particles and blurs them; `readback()`
    returns (nz, ny, nx, 4) float32 voxels of (phi in voxels, volume
    fraction, wall proximity, 1). `bounds` is ((min x, y, z), (max x, y, z));
    `wet` makes the loop's colliders wetted walls (the `colliders` option);
    `solids` is a particle range the liquid wets. `radius`,
    `meniscusHeight` and `meniscusWidth` default to 1.25, 1.6 and 2.5
    particle radii (SurfaceField.ts:101-115)."""

    def __init__(self, fluid, bounds, voxelBudget=None, motionStretch=0.0, solids=None, wet=False,
                 radius=None, meniscusHeight=None, meniscusWidth=None):
        o = _lib.SurfaceFieldOptions()
        (o.min[0], o.min[1], o.min[2]), (o.max[0], o.max[1], o.max[2]) = (
            tuple(float(v) for v in bounds[0]), tuple(float(v) for v in bounds[1]))
        o.voxel_budget = float(voxelBudget) if voxelBudget is not None else 0.0
        o.motion_stretch = float(motionStretch)
        if solids is not None:
            o.solids_start, o.solids_count = _range(solids, 0)
        o.wet = 1 if wet else 0
        o.radius = float(radius) if radius is not None else -1.0
        o.meniscus_height = float(meniscusHeight) if meniscusHeight is not None else -1.0
        o.meniscus_width = float(meniscusWidth) if meniscusWidth is not None else -1.0
        h = C.c_void_p()
        call("pf_surface_field_create", fluid.handle, C.byref(o), C.byref(h))
        self.handle = h
        self.fluid = fluid
        res = (C.c_uint32 * 3)()
        org = (C.c_double * 3)()
        vox = C.c_double()
        call("pf_surface_field_info", h, res, org, C.byref(vox))
        self.resolution = (res[0], res[1], res[2])
        self.origin = (org[0], org[1], org[2])
        self.voxel = vox.value

    def update(self):
        call("pf_surface_field_update", self.handle)

    def readback(self):
        nx, ny, nz = self.resolution
        out = np.zeros((nz, ny, nx, 4), np.float32)
        call("pf_surface_field_readback", self.handle, _ptr(out, f32p))
        return out

    def dispose(self):
        if self.handle:
            call("pf_surface_field_destroy", self.handle)
            self.handle = None


class PrimitiveSet:
    """src/core/collision/PrimitiveSet.ts:121-518."""

    def __init__(self, particles, capacity=None):
        self.particles = particles
        h = C.c_void_p()
        call("pf_primitives_create", particles.handle, int(capacity or 0), C.byref(h))
        self.handle = h

    def collider(self):
        return _lib.Collider(1, self.handle)

    def addPlane(self, normal, point, muS=None, muK=None):
        slot = C.c_uint32()
        call("pf_primitives_add_plane", self.handle, _d3(normal), _d3(point), -1.0 if muS is None else float(muS), -1.0 if muK is None else float(muK), C.byref(slot))
        return slot.value

    def addSphere(self, center, radius, invert=False, muS=None, muK=None):
        slot = C.c_uint32()
        call("pf_primitives_add_sphere", self.handle, _d3(center), float(radius), 1 if invert else 0, -1.0 if muS is None else float(muS), -1.0 if muK is None else float(muK), C.byref(slot))
        return slot.value

    def addBox(self, center, halfExtents, rotation=None, invert=False, muS=None, muK=None):
        slot = C.c_uint32()
        call("pf_primitives_add_box", self.handle, _d3(center), _d3(halfExtents), None if rotation is None else _d4(rotation), 1 if invert else 0, -1.0 if muS is None else float(muS), -1.0 if muK is None else float(muK), C.byref(slot))
        return slot.value

    def addCapsule(self, a, b, radius, muS=None, muK=None):
        slot = C.c_uint32()
        call("pf_primitives_add_capsule", self.handle, _d3(a), _d3(b), float(radius), -1.0 if muS is None else float(muS), -1.0 if muK is None else float(muK), C.byref(slot))
        return slot.value

    def setSphere(self, slot, center, radius, velocity=None):
        call("pf_primitives_set_sphere", self.handle, int(slot), _d3(center), float(radius), None if velocity is None else _d3(velocity))

    def setPose(self, slot, position, rotation=(0, 0, 0, 1), dt=0.0):
        call("pf_primitives_set_pose", self.handle, int(slot), _d3(position), _d4(rotation), float(dt))

    @property
    def count(self):
        n = C.c_uint32()
        call("pf_primitives_count", self.handle, C.byref(n))
        return n.value

    def dispose(self):
        if self.handle:
            call("pf_primitives_destroy", self.handle)
            self.handle = None


class SDFCollider:
    """src/core/collision/SDFCollider.ts:68-311.

    ``sdf`` is a dict with ``data`` (distances, x fastest, then y, then z),
    ``resolution``, ``origin`` (corner of the first voxel) and ``voxelSize``,
    as SDFData (SDFCollider.ts:32-39)."""

    def __init__(self, particles, sdf, muS=None, muK=None, position=None, rotation=None, scale=None, thickness=None):
        self.particles = particles
        data = _f32(np.asarray(sdf["data"]).reshape(-1))
        res = np.ascontiguousarray(sdf["resolution"], np.uint32)
        if res.shape != (3,) or data.size != int(res.prod()):
            raise PFError(1, f"SDFCollider: {data.size} values do not match resolution {'×'.join(str(int(r)) for r in res)}")
        o = _lib.SdfOptions()
        o.mu_s = float(muS) if muS is not None else -1.0
        o.mu_k = float(muK) if muK is not None else -1.0
        for a, v in enumerate(position if position is not None else (0.0, 0.0, 0.0)):
            o.position[a] = float(v)
        self._rot = _d4(rotation) if rotation is not None else None
        o.rotation = C.cast(self._rot, f64p) if self._rot is not None else None
        o.scale = float(scale) if scale is not None else 0.0
        o.thickness = float(thickness) if thickness is not None else 0.0
        h = C.c_void_p()
        call("pf_sdf_collider_create", particles.handle, _ptr(data, f32p), _ptr(res, u32p), _d3(sdf["origin"]),
             _d3(sdf["voxelSize"]), C.byref(o), C.byref(h))
        self.handle = h

    def collider(self):
        return _lib.Collider(2, self.handle)

    def setPosition(self, p):
        call("pf_sdf_collider_set_position", self.handle, _d3(p))

    def setRotation(self, q):
        call("pf_sdf_collider_set_rotation", self.handle, _d4(q))

    def setScale(self, s):
        call("pf_sdf_collider_set_scale", self.handle, float(s))

    def setFriction(self, muS, muK):
        call("pf_sdf_collider_set_friction", self.handle, float(muS), float(muK))

    def setThickness(self, t):
        call("pf_sdf_collider_set_thickness", self.handle, float(t))

    def dispose(self):
        if self.handle:
            call("pf_sdf_collider_destroy", self.handle)
            self.handle = None


class SoftbodySystem:
    """src/softbody/SoftbodySystem.ts:105-359.

    bodies: list of dicts with ``range`` (start, count), optional
    ``compliance``, ``surfaceCount``, ``restPositions`` ((count, 3)) and
    ``edges`` (local index pairs, required for ``shapeMatching="local"``).
    """

    def __init__(self, particles, bodies, shapeMatching="global", selfCollision=False):
        if shapeMatching not in ("global", "local"):
            raise PFError(1, "SoftbodySystem: shapeMatching must be 'global' or 'local'")
        self.particles = particles
        self.shapeMatching = shapeMatching
        defs = (_lib.SoftbodyDef * len(bodies))()
        self._keep = []
        for k, b in enumerate(bodies):
            s, c = _range(b["range"], 0)
            defs[k].start, defs[k].count = s, c
            sc = b.get("surfaceCount")
            defs[k].surface_count = 0xFFFFFFFF if sc is None else int(sc)
            defs[k].compliance = float(b.get("compliance", 0.0))
            rp = b.get("restPositions")
            if rp is not None:
                arr = _f32(rp).reshape(-1)
                self._keep.append(arr)
                defs[k].rest_positions = _ptr(arr, f32p)
                defs[k].rest_count = arr.size
            ed = b.get("edges")
            if ed is not None:
                e = np.ascontiguousarray(ed, np.uint32).reshape(-1)
                if e.size % 2:
                    raise PFError(1, f"SoftbodySystem: body {k} has an odd number of edge indices")
                self._keep.append(e)
                defs[k].edges = _ptr(e, u32p)
                defs[k].edge_count = e.size // 2
        h = C.c_void_p()
        call("pf_softbody_create", particles.handle, defs, len(bodies), 1 if shapeMatching == "local" else 0, C.byref(h))
        self.handle = h
        self.nbodies = len(bodies)
        if selfCollision:
            call("pf_softbody_set_self_collision", h, 1)

    def setCompliance(self, index, compliance):
        call("pf_softbody_set_compliance", self.handle, int(index), float(compliance))

    def readbackFrames(self):
        c = np.zeros((self.nbodies, 4), np.float32)
        r = np.zeros((self.nbodies, 3, 4), np.float32)
        call("pf_softbody_readback_frames", self.handle, _ptr(c, f32p), _ptr(r, f32p))
        return c[:, :3], r[:, :, :3]

    def material(self):
        return _lib.Material(2, self.handle)

    def dispose(self):
        if self.handle:
            call("pf_softbody_destroy", self.handle)
            self.handle = None


SimLoopOverflow = namedtuple("SimLoopOverflow", "positions velocities grid contacts neighbors impulses long_bucket")


TILE_UNION_CANDIDATES = (768, 1024, 1280, 1536, 2048)
# Launch shapes: 0 is one workgroup per tile; the others are the persistent
# form's workgroup caps, 1, 2, 4 and 8 workgroups per CU of the MI355X's 256.
LAUNCH_SHAPE_CANDIDATES = (0, 256, 512, 1024, 2048)


def _autotune(fn, handle, dt, candidates, steps):
    cand = np.ascontiguousarray(candidates, np.uint32)
    ms = np.zeros(len(cand), np.float32)
    best = C.c_uint32()
    call(fn, handle, float(dt), _ptr(cand, u32p), len(cand), int(steps), C.byref(best), _ptr(ms, f32p))
    return best.value, {int(c): float(t) for c, t in zip(cand, ms)}


class NoSelfContacts:
    """A material that only keeps its range from contacting itself: what a
    Material whose build() returns just ``noSelfContacts`` does
    (materials.ts:45, loop.ts:372-376)."""

    def __init__(self, range_):
        self.noSelfContacts = (int(range_[0]), int(range_[1]))


class ContactBuffer:
    """SimLoop.contacts (src/core/contact/ContactBuffer.ts:38-98)."""

    def __init__(self, loop, maxContacts):
        self._loop = loop
        self.maxContacts = maxContacts

    def readbackCount(self):
        """Pairs emitted in the last substep, dropped ones included."""
        n = C.c_uint32()
        call("pf_loop_readback_contacts", self._loop.handle, C.byref(n), None, 0, None)
        return n.value

    def readbackPairs(self):
        """(nContacts, pairs): the stored (i, j) pairs, flattened."""
        pairs = np.zeros(2 * self.maxContacts, np.uint32)
        n, stored = C.c_uint32(), C.c_uint32()
        call("pf_loop_readback_contacts", self._loop.handle, C.byref(n), _ptr(pairs, u32p), self.maxContacts,
             C.byref(stored))
        return stored.value, pairs[:2 * stored.value].copy()


class SimLoop:
    """src/core/loop.ts:92-324."""

    def __init__(self, particles, substeps=4, iterations=2, gravity=(0.0, -9.81, 0.0), materials=(),
                 colliders=(), hashOrigin=None, contact=None):
        self.particles = particles
        o = _lib.LoopOptions()
        self.contacts = None
        quiet = [m.noSelfContacts for m in materials if isinstance(m, NoSelfContacts)]
        materials = [m for m in materials if not isinstance(m, NoSelfContacts)]
        self._quiet = np.ascontiguousarray(np.asarray(quiet, np.uint32).reshape(-1), np.uint32)
        if contact is not None and contact is not False:
            c = {} if contact is True else dict(contact)
            o.contact = 1
            o.contact_mu_s = float(c.get("muS", -1.0))
            o.contact_mu_k = float(c.get("muK", -1.0))
            if c.get("muS") is not None and not float(c["muS"]) >= 0 or c.get("muK") is not None and not float(c["muK"]) >= 0:
                raise PFError(1, f"SimLoop: friction coefficients must be non-negative (got muS={c.get('muS')}, muK={c.get('muK')})")
            mc = c.get("maxContacts")
            if mc is not None and (int(mc) != mc or mc <= 0):
                raise PFError(1, f"ContactBuffer: maxContacts must be a positive integer, got {mc}")
            o.max_contacts = 0 if mc is None else int(mc)
            o.nquiet = len(quiet)
            o.quiet = _ptr(self._quiet, u32p) if len(quiet) else None
            self.contacts = ContactBuffer(self, int(mc) if mc is not None else 8 * particles.capacity)
        o.substeps = int(substeps)
        o.iterations = int(iterations)
        for a in range(3):
            o.gravity[a] = float(gravity[a])
        if hashOrigin is not None:
            o.has_hash_origin = 1
            for a in range(3):
                o.hash_origin[a] = float(hashOrigin[a])
        mats = (_lib.Material * max(1, len(materials)))(*[m.material() for m in materials])
        cols = (_lib.Collider * max(1, len(colliders)))(*[c.collider() for c in colliders])
        h = C.c_void_p()
        call("pf_loop_create", particles.handle, C.byref(o), mats, len(materials), cols, len(colliders), C.byref(h))
        self.handle = h
        self.materials = list(materials)
        self.colliders = list(colliders)
        self._grid = None

    def step(self, dt):
        call("pf_loop_step", self.handle, float(dt))

    def synchronize(self):
        call("pf_loop_synchronize", self.handle)

    def readbackOverflow(self):
        o = _lib.Overflow()
        call("pf_loop_readback_overflow", self.handle, C.byref(o))
        return SimLoopOverflow(*(bool(getattr(o, f)) for f in SimLoopOverflow._fields))

    def info(self):
        s, it, cell, H, k = C.c_uint32(), C.c_uint32(), C.c_double(), C.c_uint32(), C.c_uint32()
        call("pf_loop_info", self.handle, C.byref(s), C.byref(it), C.byref(cell), C.byref(H), C.byref(k))
        return {"substeps": s.value, "iterations": it.value, "cellSize": cell.value, "hashTableSize": H.value, "kernelsPerStep": k.value}

    @property
    def substeps(self):
        return self.info()["substeps"]

    @substeps.setter
    def substeps(self, value):
        call("pf_loop_set_substeps", self.handle, int(value))

    @property
    def iterations(self):
        return self.info()["iterations"]

    def setGravity(self, g):
        call("pf_loop_set_gravity", self.handle, _d3(g))

    def useGraph(self, enabled=True):
        call("pf_loop_set_graph", self.handle, 1 if enabled else 0)

    def autotune(self, dt, candidates=TILE_UNION_CANDIDATES, steps=2):
        """Time each tile-union LDS capacity on this loop's workload and keep
        the fastest; returns (best, {capacity: ms per step}). Advances the scene."""
        return _autotune("pf_loop_autotune", self.handle, dt, candidates, steps)

    def setLaunchShape(self, workgroups):
        """0: one workgroup per tile; else the persistent form, at most
        `workgroups` (a multiple of 8) workgroups per launch, each walking
        the tile list. Results do not depend on it."""
        call("pf_loop_set_launch_shape", self.handle, int(workgroups))

    def autotuneLaunch(self, dt, candidates=LAUNCH_SHAPE_CANDIDATES, steps=2):
        """Time each launch shape on this loop's workload and keep the
        fastest; returns (best, {workgroups: ms per step}). Advances the scene."""
        return _autotune("pf_loop_autotune_launch", self.handle, dt, candidates, steps)

    @property
    def hashGrid(self):
        if self._grid is None:
            h = C.c_void_p()
            call("pf_loop_grid", self.handle, C.byref(h))
            self._grid = HashGrid(self.particles, 0, _handle=h, _owner=self)
        return self._grid

    def dispose(self):
        if self.handle:
            call("pf_loop_destroy", self.handle)
            self.handle = None


class Batch:
    """Many independent SimLoops advanced by one launch sequence (no reference counterpart)."""

    def __init__(self, loops, context=None):
        ctx = context or default_context()
        arr = (C.c_void_p * len(loops))(*[L.handle for L in loops])
        h = C.c_void_p()
        call("pf_batch_create", ctx.handle, arr, len(loops), C.byref(h))
        self.handle = h
        self.loops = list(loops)

    def step(self, dt):
        call("pf_batch_step", self.handle, float(dt))

    def useGraph(self, enabled=True):
        call("pf_batch_set_graph", self.handle, 1 if enabled else 0)

    def autotune(self, dt, candidates=TILE_UNION_CANDIDATES, steps=2):
        """Time each tile-union LDS capacity on the batch's workload and keep
        the fastest; returns (best, {capacity: ms per step}). Advances the scenes."""
        return _autotune("pf_batch_autotune", self.handle, dt, candidates, steps)

    def setLaunchShape(self, workgroups):
        """0: one workgroup per tile; else the persistent form (SimLoop.setLaunchShape)."""
        call("pf_batch_set_launch_shape", self.handle, int(workgroups))

    def autotuneLaunch(self, dt, candidates=LAUNCH_SHAPE_CANDIDATES, steps=2):
        """Time each launch shape on the batch's workload and keep the
        fastest; returns (best, {workgroups: ms per step}). Advances the scenes."""
        return _autotune("pf_batch_autotune_launch", self.handle, dt, candidates, steps)

    def dispose(self):
        if self.handle:
            call("pf_batch_destroy", self.handle)
            self.handle = None


class Server:
    """The served form (no reference counterpart): step requests for live
    scenes queue as they arrive and run in multi-scene launches as slots and
    memory free up; scenes stay resident in HBM under a byte budget and idle
    scenes are evicted to host memory. Scenes need evictable particles.
    Request times are seconds on the server's clock (``now()``)."""

    STATUS = ("pending", "launched", "done", "failed")

    def __init__(self, context=None, hbmBudget=0, maxBatchScenes=0, maxBatchParticles=0, slots=0, hostBudget=0):
        ctx = context or default_context()
        o = _lib.ServerOptions(int(hbmBudget), int(maxBatchScenes), int(maxBatchParticles), int(slots),
                               int(hostBudget))
        h = C.c_void_p()
        call("pf_server_create", ctx.handle, C.byref(o), C.byref(h))
        self.handle = h
        self.context = ctx
        self.loops = {}

    def addScene(self, loop):
        sid = C.c_uint32()
        call("pf_server_add_scene", self.handle, loop.handle, C.byref(sid))
        self.loops[sid.value] = loop
        return sid.value

    def removeScene(self, sceneId):
        call("pf_server_remove_scene", self.handle, int(sceneId))
        self.loops.pop(int(sceneId), None)

    def submit(self, sceneId, dt):
        t = C.c_uint64()
        call("pf_server_submit", self.handle, int(sceneId), float(dt), C.byref(t))
        return t.value

    def wait(self, ticket):
        info = _lib.RequestInfo()
        call("pf_server_wait", self.handle, int(ticket), C.byref(info))
        return self._info(info)

    def drain(self):
        call("pf_server_drain", self.handle)

    def evictScene(self, sceneId):
        call("pf_server_evict_scene", self.handle, int(sceneId))

    def setLaunchShape(self, workgroups):
        """Launch shape of the launches formed from now on (SimLoop.setLaunchShape)."""
        call("pf_server_set_launch_shape", self.handle, int(workgroups))

    def now(self):
        return _lib.load().pf_server_now(self.handle)

    def requests(self, first, count):
        """Times of requests [first, first + count) as a structured array."""
        arr = (_lib.RequestInfo * max(1, int(count)))()
        call("pf_server_request_info", self.handle, int(first), int(count), arr)
        out = np.zeros(int(count), dtype=[("scene", np.uint32), ("status", np.uint32), ("launch_scenes", np.uint32),
                                          ("t_submit", np.float64), ("t_launch", np.float64), ("t_done", np.float64)])
        for k in range(int(count)):
            r = arr[k]
            out[k] = (r.scene, r.status, r.launch_scenes, r.t_submit, r.t_launch, r.t_done)
        return out

    def stats(self):
        st = _lib.ServerStats()
        call("pf_server_stats_get", self.handle, C.byref(st))
        return {f: getattr(st, f) for f, _ in st._fields_}

    @staticmethod
    def _info(r):
        return {"scene": r.scene, "status": Server.STATUS[r.status], "launchScenes": r.launch_scenes,
                "tSubmit": r.t_submit, "tLaunch": r.t_launch, "tDone": r.t_done}

    def dispose(self):
        if self.handle:
            call("pf_server_destroy", self.handle)
            self.handle = None


# ---------------------------------------------------------------------------
# Scene layout (src/simulation/layout.ts, Simulation.ts:957-1104)
# ---------------------------------------------------------------------------
def fill_box(box_min, box_max, spacing):
    """fillBox(box, spacing) (layout.ts:7-14): (N, 3) float32 lattice, y -> z -> x."""
    n = C.c_uint32()
    call("pf_layout_fill_box", _d3(box_min), _d3(box_max), float(spacing), None, 0, C.byref(n))
    out = np.zeros((n.value, 3), np.float32)
    call("pf_layout_fill_box", _d3(box_min), _d3(box_max), float(spacing), _ptr(out, f32p), n.value, C.byref(n))
    return out


def radius_for_budget(volume, area, cloth_spacing, budget):
    load()
    return _lib._lib.pf_layout_radius_for_budget(float(volume), float(area), float(cloth_spacing), int(budget))


def fit_budget(budget, container, boxes):
    """Simulation.fitBudget for box fluids: (radius, particle count)."""
    cont = None if container is None else (C.c_double * 6)(*[float(v) for v in container])
    flat = (C.c_double * (6 * len(boxes)))(*[float(v) for b in boxes for v in b])
    r, n = C.c_double(), C.c_uint32()
    call("pf_layout_fit_budget", int(budget), cont, flat, len(boxes), C.byref(r), C.byref(n))
    return r.value, n.value


from .simulation import Simulation  # noqa: E402
