"""ctypes binding of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker or as the timed CPU baseline.
It restates the reference's algorithm (see pf_oracle.c / orc_sim.h for the
file:line of each pass).
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")

P = C.POINTER
u32, i32, f32, f64, vp = C.c_uint32, C.c_int, C.c_float, C.c_double, C.c_void_p

_libs = {}


def build():
    subprocess.run(["make", "-s", "-C", HERE, "liboracle.so"], check=True)


def lib(path=None):
    """The oracle library (oracle/liboracle.so, built on first use), or the
    build of the oracle's sources at `path` (the planted-mistake tests load
    mutated copies that way)."""
    key = path or LIB
    if key in _libs:
        return _libs[key]
    if path is None and not os.path.exists(LIB):
        build()
    L = C.CDLL(key)
    L.orc_part1by2.argtypes = [u32]
    L.orc_part1by2.restype = u32
    L.orc_morton.argtypes = [C.c_int32, C.c_int32, C.c_int32]
    L.orc_morton.restype = u32
    L.orc_grid_build.argtypes = [P(f32), u32, f32, P(f32), u32, P(u32), P(u32), P(u32), P(u32), P(u32), P(u32)]
    L.orc_grid_build.restype = i32
    L.orc_grid_query_counts.argtypes = [P(f32), u32, f32, P(f32), u32, P(f32), u32, f32, P(u32), P(u32)]
    L.orc_grid_query_counts.restype = None
    L.orc_float_to_half.argtypes = [f32]
    L.orc_float_to_half.restype = C.c_uint16
    L.orc_half_to_double.argtypes = [C.c_uint16]
    L.orc_half_to_double.restype = f64
    for suf in ("f32", "f64"):
        g = lambda n: getattr(L, f"orc_{n}_{suf}")
        g("scene_create").argtypes = [u32, f64]
        g("scene_create").restype = vp
        g("scene_destroy").argtypes = [vp]
        g("scene_destroy").restype = None
        g("scene_set_particles").argtypes = [vp, P(f32), P(f32), P(f32)]
        g("scene_set_particles").restype = None
        g("scene_set_loop").argtypes = [vp, u32, u32, P(f64), f64, u32, P(f64)]
        g("scene_set_loop").restype = None
        g("scene_add_fluid").argtypes = [vp, u32, u32, f64, f64, f64, f64, u32, f64, f64, f64, f64]
        g("scene_add_fluid").restype = i32
        g("scene_add_prim").argtypes = [vp, u32, u32, P(f64), P(f64), P(f64), P(f64), P(f64)]
        g("scene_add_prim").restype = i32
        g("scene_add_viscosity").argtypes = [vp, u32, f64, u32]
        g("scene_add_viscosity").restype = i32
        g("scene_set_prim").argtypes = [vp, u32, u32, u32, P(f64), P(f64), P(f64), P(f64), P(f64)]
        g("scene_set_prim").restype = i32
        g("scene_add_body").argtypes = [vp, u32, u32, f64, P(f32), P(f32)]
        g("scene_add_body").restype = i32
        g("scene_add_sdf").argtypes = [vp, P(f32), P(u32), P(f64), P(f64), P(f64), P(f64), f64, f64, f64, f64]
        g("scene_add_sdf").restype = i32
        g("scene_sdf_sample").argtypes = [vp, u32, P(f64), P(f64), P(f64)]
        g("scene_sdf_sample").restype = None
        g("scene_add_body_edges").argtypes = [vp, P(u32), u32]
        g("scene_add_body_edges").restype = i32
        g("scene_get_orientation").argtypes = [vp, P(f64), P(f64), P(f64)]
        g("scene_get_orientation").restype = None
        g("scene_add_boundary").argtypes = [vp, u32, u32, u32, i32]
        g("scene_add_boundary").restype = i32
        g("scene_get_bvol").argtypes = [vp, P(f64)]
        g("scene_get_bvol").restype = None
        g("scene_set_groups").argtypes = [vp, P(u32)]
        g("scene_set_groups").restype = None
        g("scene_set_contact").argtypes = [vp, f64, f64, u32, P(u32), u32]
        g("scene_set_contact").restype = i32
        g("scene_contacts").argtypes = [vp, P(u32), u32]
        g("scene_contacts").restype = u32
        g("scene_surface_field").argtypes = [vp, P(f64), P(f64), f64, f64, f64, P(u32), i32, P(f64), vp, P(u32), P(f64)]
        g("scene_surface_field").restype = i32
        g("scene_step").argtypes = [vp, f64]
        g("scene_step").restype = None
        g("scene_get").argtypes = [vp, P(f64), P(f64), P(f64), P(f64)]
        g("scene_get").restype = None
        g("scene_flags").argtypes = [vp, P(i32)]
        g("scene_flags").restype = None
        g("scene_body_frame").argtypes = [vp, u32, P(f64), P(f64)]
        g("scene_body_frame").restype = None
        g("polar").argtypes = [P(f64), P(f64)]
        g("polar").restype = None
        g("scene_neighbors").argtypes = [vp, u32, P(u32), P(u32)]
        g("scene_neighbors").restype = None
        g("scene_fluid_fields").argtypes = [vp, u32] + [P(f64)] * 7
        g("scene_fluid_fields").restype = None
    L.orc_set_threads.argtypes = [i32]
    L.orc_set_threads.restype = None
    L.orc_get_threads.argtypes = []
    L.orc_get_threads.restype = i32
    L.orc_poly6_f32.argtypes = [f64, f64]
    L.orc_poly6_f32.restype = f64
    L.orc_spiky_f32.argtypes = [P(f64), f64, P(f64)]
    L.orc_spiky_f32.restype = None
    _libs[key] = L
    return L


def set_threads(n):
    """Threads of the oracle's particle loops (0: all the host's cores); the
    results do not depend on it."""
    lib().orc_set_threads(int(n))


def get_threads():
    return int(lib().orc_get_threads())


def _p(a, t):
    return None if a is None else a.ctypes.data_as(P(t))


def morton(cx, cy, cz):
    return lib().orc_morton(int(cx), int(cy), int(cz))


def part1by2(n):
    return lib().orc_part1by2(int(n))


def grid_build(pos4, cell, H, origin=(0.0, 0.0, 0.0)):
    """Stable counting sort: dict of cellIndex, counts, cellStart, cellEnd,
    sortedIndices, sortedKeys, overflow (counts etc. padded to 1024)."""
    pos = np.ascontiguousarray(pos4, np.float32).reshape(-1, 4)
    n = pos.shape[0]
    Hp = (H + 1023) // 1024 * 1024
    out = dict(cellIndex=np.zeros(n, np.uint32), counts=np.zeros(Hp, np.uint32), cellStart=np.zeros(Hp, np.uint32),
               cellEnd=np.zeros(Hp, np.uint32), sortedIndices=np.zeros(n, np.uint32), sortedKeys=np.zeros(n, np.uint32))
    o = np.asarray(origin, np.float32)
    over = lib().orc_grid_build(_p(pos, f32), n, f32(cell), _p(o, f32), H, _p(out["cellIndex"], u32), _p(out["counts"], u32),
                                _p(out["cellStart"], u32), _p(out["cellEnd"], u32), _p(out["sortedIndices"], u32),
                                _p(out["sortedKeys"], u32))
    out["overflow"] = bool(over)
    return out


def grid_query_counts(pos4, cell, H, query4, radius, origin=(0.0, 0.0, 0.0)):
    pos = np.ascontiguousarray(pos4, np.float32).reshape(-1, 4)
    q = np.ascontiguousarray(query4, np.float32).reshape(-1, 4)
    o = np.asarray(origin, np.float32)
    cand = np.zeros(q.shape[0], np.uint32)
    within = np.zeros(q.shape[0], np.uint32)
    lib().orc_grid_query_counts(_p(pos, f32), pos.shape[0], f32(cell), _p(o, f32), H, _p(q, f32), q.shape[0], f32(radius),
                                _p(cand, u32), _p(within, u32))
    return cand, within


def float_to_half(v):
    """Half-float bits of v as three's DataUtils.toHalfFloat gives them."""
    return int(lib().orc_float_to_half(float(v)))


def half_to_float(h):
    return float(lib().orc_half_to_double(int(h)))


def polar(a, precision="f32"):
    a9 = np.ascontiguousarray(np.asarray(a, np.float64).reshape(9))
    r9 = np.zeros(9, np.float64)
    getattr(lib(), f"orc_polar_{precision}")(_p(a9, f64), _p(r9, f64))
    return r9.reshape(3, 3)


class Scene:
    """One scene stepped on the CPU, in FP32 ("f32") or FP64 ("f64")."""

    def __init__(self, n, radius, precision="f32", library=None):
        self.suf = precision
        self.radius = radius
        self.L = lib() if library is None else library
        self.n = int(n)
        self.h = self._f("scene_create")(self.n, float(radius))

    def _f(self, name):
        return getattr(self.L, f"orc_{name}_{self.suf}")

    def set_particles(self, x3, v3=None, w=None):
        x = np.ascontiguousarray(x3, np.float32).reshape(-1)
        v = None if v3 is None else np.ascontiguousarray(v3, np.float32).reshape(-1)
        ww = None if w is None else np.ascontiguousarray(w, np.float32)
        self._f("scene_set_particles")(self.h, _p(x, f32), _p(v, f32), _p(ww, f32))

    def set_loop(self, substeps, iterations, gravity, cell, H, origin=(0.0, 0.0, 0.0)):
        g = np.asarray(gravity, np.float64)
        o = np.asarray(origin, np.float64)
        self._f("scene_set_loop")(self.h, int(substeps), int(iterations), _p(g, f64), float(cell), int(H), _p(o, f64))

    def add_fluid(self, start, count, rho0, spacing, h, compliance=1e-4, viscosity=None, vorticity=None,
                  surface_tension=None, adhesion=None):
        eff = 0
        eff |= 1 if viscosity is not None else 0
        eff |= 2 if vorticity is not None else 0
        eff |= 4 if surface_tension is not None else 0
        eff |= 8 if adhesion is not None else 0
        rc = self._f("scene_add_fluid")(self.h, int(start), int(count), float(rho0), float(spacing), float(h),
                                        float(compliance), eff, float(viscosity or 0), float(vorticity or 0),
                                        float(surface_tension or 0), float(adhesion or 0))
        assert rc == 0

    def add_viscosity(self, fluid, viscosity, iterations=12):
        """ViscositySolver after fluid `fluid` (ViscositySolver.ts:49-123)."""
        rc = self._f("scene_add_viscosity")(self.h, int(fluid), float(viscosity), int(iterations))
        assert rc == 0

    def add_prim(self, kind, flags, d0, d1, vel, rot=(0, 0, 0, 1), spin=(0, 0, 0)):
        arr = [np.asarray(a, np.float64) for a in (d0, d1, vel, rot, spin)]
        rc = self._f("scene_add_prim")(self.h, int(kind), int(flags), *[_p(a, f64) for a in arr])
        assert rc == 0

    def set_prim(self, k, kind, flags, d0, d1, vel, rot=(0, 0, 0, 1), spin=(0, 0, 0)):
        arr = [np.asarray(a, np.float64) for a in (d0, d1, vel, rot, spin)]
        rc = self._f("scene_set_prim")(self.h, int(k), int(kind), int(flags), *[_p(a, f64) for a in arr])
        assert rc == 0

    # PrimitiveSet add* (PrimitiveSet.ts:160-209): d0/d1 hold the shape, d1.w
    # static friction, vel.w kinetic friction; flags bit 0 inverts the shape.
    def add_sphere(self, center, radius, invert=False, mus=0.5, muk=0.4):
        c = np.asarray(center, np.float64)
        self.add_prim(1, 1 if invert else 0, [c[0], c[1], c[2], radius], [0, 0, 0, mus], [0, 0, 0, muk])

    def add_box(self, center, half_extents, rotation=None, invert=False, mus=0.5, muk=0.4):
        c, he = np.asarray(center, np.float64), np.asarray(half_extents, np.float64)
        q = np.array([0, 0, 0, 1.0]) if rotation is None else np.asarray(rotation, np.float64)
        q = q / np.sqrt((q * q).sum())
        self.add_prim(2, 1 if invert else 0, [c[0], c[1], c[2], 0], [he[0], he[1], he[2], mus], [0, 0, 0, muk], q)

    def add_capsule(self, a, b, radius, mus=0.5, muk=0.4):
        a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
        self.add_prim(3, 0, [a[0], a[1], a[2], radius], [b[0], b[1], b[2], mus], [0, 0, 0, muk])

    def add_plane(self, normal, point, mus=0.5, muk=0.4):
        n = np.asarray(normal, np.float64)
        n = n * (1.0 / np.sqrt((n * n).sum()))
        self.add_prim(0, 0, [n[0], n[1], n[2], 0], [point[0], point[1], point[2], mus], [0, 0, 0, muk])

    def add_body(self, start, count, compliance, rest3, weights):
        r = np.ascontiguousarray(rest3, np.float32).reshape(-1)
        w = np.ascontiguousarray(weights, np.float32)
        rc = self._f("scene_add_body")(self.h, int(start), int(count), float(compliance), _p(r, f32), _p(w, f32))
        assert rc == 0

    def add_sdf(self, data, resolution, origin, voxel_size, position=(0.0, 0.0, 0.0), rotation=None, scale=1.0,
                thickness=0.0, mus=0.5, muk=0.4):
        """An SDF collider; rotation is a row-major 3x3 matrix (identity if None)."""
        d = np.ascontiguousarray(data, np.float32).reshape(-1)
        res = np.asarray(resolution, np.uint32)
        rot = np.eye(3) if rotation is None else np.asarray(rotation, np.float64).reshape(3, 3)
        arr = [np.ascontiguousarray(a, np.float64).reshape(-1) for a in (origin, voxel_size, position, rot)]
        rc = self._f("scene_add_sdf")(self.h, _p(d, f32), _p(res, u32), *[_p(a, f64) for a in arr], float(scale),
                                      float(thickness), float(mus), float(muk))
        assert rc == 0

    def sdf_sample(self, k, x):
        xx = np.asarray(x, np.float64)
        phi = np.zeros(1)
        g = np.zeros(3)
        self._f("scene_sdf_sample")(self.h, int(k), _p(xx, f64), _p(phi, f64), _p(g, f64))
        return float(phi[0]), g

    def add_body_edges(self, edges):
        """Edges (local pairs) of the last added body; makes shape matching local."""
        e = np.ascontiguousarray(edges, np.uint32).reshape(-1)
        rc = self._f("scene_add_body_edges")(self.h, _p(e, u32), len(e) // 2)
        assert rc == 0

    def orientation(self):
        q, qp, om = (np.zeros((self.n, 4)) for _ in range(3))
        self._f("scene_get_orientation")(self.h, _p(q, f64), _p(qp, f64), _p(om, f64))
        return q, qp, om

    def add_boundary(self, fluid, start, count, dynamic=True):
        rc = self._f("scene_add_boundary")(self.h, int(fluid), int(start), int(count), 1 if dynamic else 0)
        assert rc == 0

    def boundary_volumes(self):
        out = np.zeros(self.n)
        self._f("scene_get_bvol")(self.h, _p(out, f64))
        return out

    def set_groups(self, groups):
        g = np.ascontiguousarray(groups, np.uint32)
        assert g.size == self.n
        self._f("scene_set_groups")(self.h, _p(g, u32))

    def set_contact(self, mus=0.5, muk=0.4, max_contacts=None, quiet=()):
        """SimLoop `contact` option (loop.ts:23-31); quiet: extra (start, count)
        ranges that never contact each other (fluids are always quiet)."""
        q = np.ascontiguousarray(np.asarray(quiet, np.uint32).reshape(-1), np.uint32)
        mc = 8 * self.n if max_contacts is None else int(max_contacts)
        rc = self._f("scene_set_contact")(self.h, float(mus), float(muk), mc, _p(q, u32), len(q) // 2)
        assert rc == 0
        self.ct_max = mc

    def contacts(self):
        """(emitted count, stored (i, j) pairs) of the last substep."""
        cap = self.ct_max
        pairs = np.zeros((cap, 2), np.uint32)
        count = int(self._f("scene_contacts")(self.h, _p(pairs, u32), cap))
        return count, pairs[:min(count, cap)].copy()

    def surface_field(self, mn, mx, voxel_budget, spacing, fluid_range, solid_range=(0, 0), stretch=0.0, wet=False,
                      radius=None, meniscus_height=None, meniscus_width=None):
        """One frame of SurfaceField (SurfaceField.ts:101-369) over the
        committed positions: (field (nz, ny, nx, 4) of phi in voxels, volume
        fraction, wall proximity, 1; voxel size)."""
        r = float(self.radius)
        uni = np.array([r * 1.25 if radius is None else radius, r * 1.6 if meniscus_height is None else meniscus_height,
                        r * 2.5 if meniscus_width is None else meniscus_width], np.float64)
        lo, hi = np.asarray(mn, np.float64), np.asarray(mx, np.float64)
        rng = np.array([fluid_range[0], fluid_range[1], solid_range[0], solid_range[1]], np.uint32)
        res = np.zeros(3, np.uint32)
        vox = np.zeros(1)
        f = self._f("scene_surface_field")
        assert f(self.h, _p(lo, f64), _p(hi, f64), float(voxel_budget), float(stretch), float(spacing), _p(rng, u32),
                 1 if wet else 0, _p(uni, f64), None, _p(res, u32), _p(vox, f64)) == 0
        dt = np.float32 if self.suf == "f32" else np.float64
        out = np.zeros((int(res[2]), int(res[1]), int(res[0]), 4), dt)
        assert f(self.h, _p(lo, f64), _p(hi, f64), float(voxel_budget), float(stretch), float(spacing), _p(rng, u32),
                 1 if wet else 0, _p(uni, f64), out.ctypes.data_as(C.c_void_p), _p(res, u32), _p(vox, f64)) == 0
        return out, float(vox[0])

    def step(self, dt):
        self._f("scene_step")(self.h, float(dt))

    def get(self):
        x = np.zeros((self.n, 3))
        xs = np.zeros((self.n, 3))
        v = np.zeros((self.n, 3))
        rho = np.zeros(self.n)
        self._f("scene_get")(self.h, _p(x, f64), _p(xs, f64), _p(v, f64), _p(rho, f64))
        return dict(x=x, xs=xs, v=v, rho=rho)

    def flags(self):
        f = np.zeros(5, np.int32)
        self._f("scene_flags")(self.h, _p(f, i32))
        return dict(grid=bool(f[0]), neighbors=bool(f[1]), positions=bool(f[2]), velocities=bool(f[3]), impulses=bool(f[4]))

    def body_frame(self, b):
        c = np.zeros(3)
        r = np.zeros(9)
        self._f("scene_body_frame")(self.h, int(b), _p(c, f64), _p(r, f64))
        return c, r.reshape(3, 3)

    def neighbors(self, f=0):
        nl = np.zeros((self.n, 64), np.uint32)
        cnt = np.zeros(self.n, np.uint32)
        self._f("scene_neighbors")(self.h, int(f), _p(nl, u32), _p(cnt, u32))
        return nl, cnt

    def fluid_fields(self, f=0):
        """Fluid f's per-pass fields of the last substep, by particle id:
        lambda and delta x of the last iteration, colour normal, omega,
        |omega|, eta and the XSPH velocity change."""
        n = self.n
        out = dict(lam=np.zeros(n), dx=np.zeros((n, 3)), normal=np.zeros((n, 3)), omega=np.zeros((n, 3)),
                   omlen=np.zeros(n), eta=np.zeros((n, 3)), dv=np.zeros((n, 3)))
        self._f("scene_fluid_fields")(self.h, int(f), *[_p(out[k], f64) for k in
                                                        ("lam", "dx", "normal", "omega", "omlen", "eta", "dv")])
        return out

    def __del__(self):
        try:
            if self.h:
                self._f("scene_destroy")(self.h)
                self.h = None
        except Exception:
            pass
