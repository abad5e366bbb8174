// Node N-API addon over the C ABI of libpf.so (include/pf.h).
//
// Thin by design: every export converts JS values to plain pointers and
// sizes, calls one pf_* entry point, and throws a JS Error carrying
// pf_last_error() on failure, so index.js can mirror the reference's classes.
// Handles cross as napi externals. loopStepAsync enqueues the step on the
// library's stream from the JS thread and resolves its Promise from a worker
// thread that waits on the stream (reference: SimLoop.step returns a Promise,
// src/core/loop.ts:267-279).
#include <node_api.h>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/pf.h"

namespace {

#define CHECK_NAPI(expr)                                                     \
  do {                                                                       \
    if ((expr) != napi_ok) {                                                 \
      napi_throw_error(env, nullptr, "pf addon: N-API call failed: " #expr); \
      return nullptr;                                                        \
    }                                                                        \
  } while (0)

bool fail_if(napi_env env, int rc) {
  if (rc == PF_OK) return false;
  napi_throw_error(env, nullptr, pf_last_error());
  return true;
}

#define PF_CALL(expr)                 \
  do {                                \
    if (fail_if(env, (expr))) return nullptr; \
  } while (0)

struct Args {
  napi_value argv[8];
  size_t argc = 8;
};

Args get_args(napi_env env, napi_callback_info info) {
  Args a;
  napi_get_cb_info(env, info, &a.argc, a.argv, nullptr, nullptr);
  for (size_t i = a.argc; i < 8; ++i) napi_get_undefined(env, &a.argv[i]);
  return a;
}

bool is_nullish(napi_env env, napi_value v) {
  napi_valuetype t;
  napi_typeof(env, v, &t);
  return t == napi_undefined || t == napi_null;
}

void* handle(napi_env env, napi_value v) {
  if (is_nullish(env, v)) return nullptr;
  void* p = nullptr;
  napi_get_value_external(env, v, &p);
  return p;
}

napi_value make_handle(napi_env env, void* p) {
  napi_value v;
  napi_create_external(env, p, nullptr, nullptr, &v);
  return v;
}

double num(napi_env env, napi_value v, double dflt = 0) {
  if (is_nullish(env, v)) return dflt;
  double d = dflt;
  napi_get_value_double(env, v, &d);
  return d;
}

uint32_t u32(napi_env env, napi_value v, uint32_t dflt = 0) {
  if (is_nullish(env, v)) return dflt;
  double d = dflt;
  napi_get_value_double(env, v, &d);
  return (uint32_t)d;
}

// The TypedArray kind each element type is read from.
template <typename T>
struct TypedKind;
template <>
struct TypedKind<float> {
  static bool ok(napi_typedarray_type t) { return t == napi_float32_array; }
  static constexpr const char* name = "a Float32Array";
};
template <>
struct TypedKind<uint32_t> {
  static bool ok(napi_typedarray_type t) { return t == napi_uint32_array; }
  static constexpr const char* name = "a Uint32Array";
};
template <>
struct TypedKind<uint8_t> {
  static bool ok(napi_typedarray_type t) { return t == napi_uint8_array || t == napi_uint8_clamped_array; }
  static constexpr const char* name = "a Uint8Array";
};

// Pointer into a TypedArray (null for null/undefined); *count = element count.
// Any other value, or a TypedArray of another element type, throws a
// TypeError and returns null (check pending() before using the result).
template <typename T>
T* typed(napi_env env, napi_value v, size_t* count = nullptr) {
  if (count) *count = 0;
  if (is_nullish(env, v)) return nullptr;
  napi_typedarray_type type;
  size_t length = 0;
  void* data = nullptr;
  napi_value buf;
  size_t offset = 0;
  bool is_typed = false;
  napi_is_typedarray(env, v, &is_typed);
  if (!is_typed || napi_get_typedarray_info(env, v, &type, &length, &data, &buf, &offset) != napi_ok ||
      !TypedKind<T>::ok(type)) {
    std::string msg = std::string("pf addon: expected ") + TypedKind<T>::name;
    napi_throw_type_error(env, nullptr, msg.c_str());
    return nullptr;
  }
  if (count) *count = length;
  return static_cast<T*>(data);
}

bool pending(napi_env env) {
  bool p = false;
  napi_is_exception_pending(env, &p);
  return p;
}

napi_value throw_range(napi_env env, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  napi_throw_range_error(env, nullptr, buf);
  return nullptr;
}

// A JS array of numbers into doubles.
std::vector<double> doubles(napi_env env, napi_value v) {
  std::vector<double> out;
  if (is_nullish(env, v)) return out;
  uint32_t n = 0;
  napi_get_array_length(env, v, &n);
  for (uint32_t i = 0; i < n; ++i) {
    napi_value e;
    napi_get_element(env, v, i, &e);
    out.push_back(num(env, e));
  }
  return out;
}

napi_value f32_array(napi_env env, size_t n, float** data) {
  napi_value ab, arr;
  void* p = nullptr;
  napi_create_arraybuffer(env, n * sizeof(float), &p, &ab);
  napi_create_typedarray(env, napi_float32_array, n, ab, 0, &arr);
  *data = static_cast<float*>(p);
  return arr;
}

napi_value u32_array(napi_env env, size_t n, uint32_t** data) {
  napi_value ab, arr;
  void* p = nullptr;
  napi_create_arraybuffer(env, n * sizeof(uint32_t), &p, &ab);
  napi_create_typedarray(env, napi_uint32_array, n, ab, 0, &arr);
  *data = static_cast<uint32_t*>(p);
  return arr;
}

napi_value number(napi_env env, double d) {
  napi_value v;
  napi_create_double(env, d, &v);
  return v;
}

napi_value boolean(napi_env env, bool b) {
  napi_value v;
  napi_get_boolean(env, b, &v);
  return v;
}

void set(napi_env env, napi_value obj, const char* key, napi_value v) { napi_set_named_property(env, obj, key, v); }

napi_value prop(napi_env env, napi_value obj, const char* key) {
  napi_value v;
  bool has = false;
  napi_has_named_property(env, obj, key, &has);
  if (!has) {
    napi_get_undefined(env, &v);
    return v;
  }
  napi_get_named_property(env, obj, key, &v);
  return v;
}

// ---------------------------------------------------------------- context
napi_value ContextCreate(napi_env env, napi_callback_info info) {
  Args a = get_args(env, info);
  pf_context* c = nullptr;
  PF_CALL(pf_context_create((int)num(env, a.argv[0]), &c));
  return make_handle(env, c);
}
napi_value ContextDestroy(napi_env env, napi_callback_info info) {
  Args a = get_args(env, info);
  PF_CALL(pf_context_destroy((pf_context*)handle(env, a.argv[0])));
  return nullptr;
}
napi_value ContextSynchronize(napi_env env, napi_callback_info info) {
  Args a = get_args(env, info);
  PF_CALL(pf_context_synchronize((pf_context*)handle(env, a.argv[0])));
  return nullptr;
}

// ---------------------------------------------------------------- particles
// (ctx, capacity, radius, evictable)
napi_value ParticlesCreate(napi_env env, napi_callback_info info) {
  Args a = get_args(env, info);
  pf_particles* p = nullptr;
  bool evictable = false;
  if (!is_nullish(env, a.argv[3])) napi_get_value_bool(env, a.argv[3], &evictable);
  auto* ctx = (pf_context*)handle(env, a.argv[0]);
  if (evictable) PF_CALL(pf_particles_create_evictable(ctx, u32(env, a.argv[1]), num(env, a.argv[2]), &p));
  else PF_CALL(pf_particles_create(ctx, u32(env, a.argv[1]), num(env, a.argv[2]), &p));
  return make_handle(env, p);
}
napi_value ParticlesDeviceBytes(napi_env env, napi_callback_info info) {
  Args a = get_args(env, info);
  uint64_t b = 0;
  PF_CALL(pf_particles_device_bytes((pf_particles*)handle(env, a.argv[0]), &b));
  return number(env, (double)b);
}
napi_value ParticlesDestroy(napi_env env, napi_callback_info info) {
  Args a = get_args(env, info);
  PF_CALL(pf_particles_destroy((pf_particles*)handle(env, a.argv[0])));
  return nullptr;
}
// (ps, start, positions3, velocities3|null, invMass|null, groups|null)
napi_value ParticlesUpload(napi_env env, napi_callback_info info) {
  Args a = get_args(env, info);
  size_t n3 = 0, nv = 0, nm = 0, ng = 0;
  float* pos = typed<float>(env, a.argv[2], &n3);
  float* vel = pending(env) ? nullptr : typed<float>(env, a.argv[3], &nv);
  float* inv = pending(env) ? nullptr : typed<float>(env, a.argv[4], &nm);
  uint32_t* grp = pending(env) ? nullptr : typed<uint32_t>(env, a.argv[5], &ng);
  if (pending(env)) return nullptr;
  // Every array covers the same particles; the library reads count entries of each.
  if (!pos || n3 % 3 != 0)
    return throw_range(env, "ParticleSystem.uploadParticles: positions must hold xyz triples, got %zu values", n3);
  const size_t n = n3 / 3;
  if (vel && nv != 3 * n)
    return throw_range(env, "ParticleSystem.uploadParticles: %zu velocity values for %zu particles", nv, n);
  if (inv && nm != n)
    return throw_range(env, "ParticleSystem.uploadParticles: %zu inverse masses for %zu particles", nm, n);
  if (grp && ng != n)
    return throw_range(env, "ParticleSystem.uploadParticles: %zu collision groups for %zu particles", ng, n);
  PF_CALL(pf_particles_upload((pf_particles*)handle(env, a.argv[0]), u32(env, a.argv[1]), (uint32_t)n, pos, vel,
                              inv, grp));
  return nullptr;
}
napi_value ParticlesSetInvMass(napi_env env, napi_callback_info info) {
  Args a = get_args(env, info);
  PF_CALL(pf_particles_set_inv_mass((pf_particles*)handle(env, a.argv[0]), u32(env, a.argv[1]), u32(env, a.argv[2]),
                                    num(env, a.argv[3])));
  return nullptr;
}
napi_value ParticlesSetCollisionGroup(napi_env env, napi_callback_info info) {
  Args a = get_args(env, info);
  PF_CALL(pf_particles_set_collision_group((pf_particles*)handle(env, a.argv[0]), u32(env, a.argv[1]),
                                           u32(env, a.argv[2]), u32(env, a.argv[3])));
  return nullptr;
}
napi_value ParticlesReadback(napi_env env, napi_callback_info info) {
  Args a = get_args(env, info);
  pf_particles* ps = (pf_particles*)handle(env, a.argv[0]);
  uint32_t cap = 0;
  double radius = 0;
  PF_CALL(pf_particles_info(ps, &cap, &radius));
  pf_snapshot s{};
  napi_value obj;
  napi_create_object(env, &obj);
  set(env, obj, "capacity", number(env, cap));
  set(env, obj, "positions", f32_array(env, 4 * (size_t)cap, &s.positions));
  set(env, obj, "predictedPositions", f32_array(env, 4 * (size_t)cap, &s.predicted_positions));
  set(env, obj, "velocities", f32_array(env, 4 * (size_t)cap, &s.velocities));
  set(env, obj, "invMass", f32_array(env, cap, &s.inv_mass));
  set(env, obj, "rotation", f32_array(env, 4 * (size_t)cap, &s.rotation));
  set(env, obj, "predictedRotation", f32_array(env, 4 * (size_t)cap, &s.predicted_rotation));
  set(env, obj, "angularVelocity", f32_array(env, 4 * (size_t)cap, &s.angular_velocity));
  PF_CALL(pf_particles_readback(ps, &s));
  return obj;
}

// ---------------------------------------------------------------- grid
napi_value GridCreate(napi_env env, napi_callback_info info) {
  Args a = get_args(env, info);
  std::vector<double> o = doubles(env, a.argv[3]);
  pf_grid* g = nullptr;
  PF_CALL(pf_grid_create((pf_particles*)handle(env, a.argv[0]), num(env, a.argv[1]), u32(env, a.argv[2]),
                         o.size() == 3 ? o.data() : nullptr, &g));
  return make_handle(env, g);
}
napi_value GridRebuild(napi_env env, napi_callback_info info) {
  Args a = get_args(env, info);
  PF_CALL(pf_grid_rebuild((pf_grid*)handle(env, a.argv[0])));
  return nullptr;
}
napi_value GridReadback(napi_env env, napi_callback_info info) {
  Args a = get_args(env, info);
  pf_grid* g = (pf_grid*)handle(env, a.argv[0]);
  uint32_t cap = 0, H = 0, Hp = 0;
  PF_CALL(pf_grid_info(g, &cap, &H, &Hp));
  uint32_t *ci, *cn, *cs, *ce, *si;
  napi_value obj;
  napi_create_object(env, &obj);
  set(env, obj, "capacity", number(env, cap));
  set(env, obj, "hashTableSize", number(env, H));
  set(env, obj, "hashTableSizePadded", number(env, Hp));
  set(env, obj, "cellIndex", u32_array(env, cap, &ci));
  set(env, obj, "counts", u32_array(env, Hp, &cn));
  set(env, obj, "cellStart", u32_array(env, Hp, &cs));
  set(env, obj, "cellEnd", u32_array(env, Hp, &ce));
  set(env, obj, "sortedIndices", u32_array(env, cap, &si));
  PF_CALL(pf_grid_readback(g, ci, cn, cs, ce, si));
  return obj;
}
napi_value GridReadbackOverflow(napi_env env, napi_callback_info info) {
  Args a = get_args(env, info);
  int o = 0;
  PF_CALL(pf_grid_readback_overflow((pf_grid*)handle(env, a.argv[0]), &o));
  return boolean(env, o != 0);
}
napi_value GridInfo(napi_env env, napi_callback_info info) {
  Args a = get_args(env, info);
  uint32_t cap = 0, H = 0, Hp = 0;
  PF_CALL(pf_grid_info((pf_grid*)handle(env, a.argv[0]), &cap, &H, &Hp));
  napi_value obj;
  napi_create_object(env, &obj);
  set(env, obj, "capacity", number(env, cap));
  set(env, obj, "hashTableSize", number(env, H));
  set(env, obj, "hashTableSizePadded", number(env, Hp));
  return obj;
}
napi_value GridSetOrigin(napi_env env, napi_callback_info info) {
  Args a = get_args(env, info);
  std::vector<double> o = doubles(env, a.argv[1]);
  if (o.size() != 3) {
    napi_throw_type_error(env, nullptr, "HashGrid: hashOrigin needs three numbers");
    return nullptr;
  }
  PF_CALL(pf_grid_set_origin((pf_grid*)handle(env, a.argv[0]), o.data()));
  return nullptr;
}
napi_value GridDestroy(napi_env env, napi_callback_info info) {
  Args a = get_args(env, info);
  PF_CALL(pf_grid_destroy((pf_grid*)handle(env, a.argv[0])));
  return nullptr;
}

// ---------------------------------------------------------------- fluid
napi_value FluidCreate(napi_env env, napi_callback_info info) {
  Args a = get_args(env, info);
  napi_value o = a.argv[1];
  pf_fluid_options opt{};
  opt.range_start = u32(env, prop(env, o, "rangeStart"));
  opt.range_count = u32(env, prop(env, o, "rangeCount"));
  opt.rest_density = num(env, prop(env, o, "restDensity"), 0);
  opt.particle_spacing = num(env, prop(env, o, "particleSpacing"), 0);
  opt.smoothing_radius = num(env, prop(env, o, "smoothingRadius"), 0);
  opt.compliance = num(env, prop(env, o, "compliance"), -1);
  const char* keys[4] = {"viscosity", "vorticity", "surfaceTension", "adhesion"};
  double* vals[4] = {&opt.viscosity, &opt.vorticity, &opt.surface_tension, &opt.adhesion};
  for (int k = 0; k < 4; ++k) {
    napi_value v = prop(env, o, keys[k]);
    if (!is_nullish(env, v)) {
      opt.present |= 1u << k;
      *vals[k] = num(env, v);
    }
  }
  pf_fluid* f = nullptr;
  PF_CALL(pf_fluid_create((pf_particles*)handle(env, a.argv[0]), &opt, &f));
  return make_handle(env, f);
}
napi_value FluidSet(napi_env env, napi_callback_info info) {
  Args a = get_args(env, info);
  PF_CALL(pf_fluid_set((pf_fluid*)handle(env, a.argv[0]), u32(env, a.argv[1]), num(env, a.argv[2])));
  return nullptr;
}
napi_value FluidGet(napi_env env, napi_callback_info info) {
  Args a = get_args(env, info);
  double v = 0;
  PF_CALL(pf_fluid_get((pf_fluid*)handle(env, a.argv[0]), u32(env, a.argv[1]), &v));
  return number(env, v);
}
napi_value FluidAddBoundary(napi_env env, napi_callback_info info) {
  Args a = get_args(env, info);
  bool dyn = true;
  if (!is_nullish(env, a.argv[3])) napi_get_value_bool(env, a.argv[3], &dyn);
  PF_CALL(pf_fluid_add_boundary((pf_fluid*)handle(env, a.argv[0]), u32(env, a.argv[1]), u32(env, a.argv[2]), dyn ? 1 : 0));
  return nullptr;
}
napi_value FluidReadbackDensity(napi_env env, napi_callback_info info) {
  Args a = get_args(env, info);
  pf_fluid* f = (pf_fluid*)handle(env, a.argv[0]);
  uint32_t cap = u32(env, a.argv[1]);
  float* d;
  napi_value arr = f32_array(env, cap, &d);
  PF_CALL(pf_fluid_readback_density(f, d));
  return arr;
}
napi_value FluidReadbackOverflow(napi_env env, napi_callback_info info) {
  Args a = get_args(env, info);
  int o = 0;
  PF_CALL(pf_fluid_readback_overflow((pf_fluid*)handle(env, a.argv[0]), &o));
  return boolean(env, o != 0);
}
napi_value FluidInfo(napi_env env, napi_callback_info info) {
  Args a = get_args(env, info);
  double m, h, sp, rho;
  PF_CALL(pf_fluid_info((pf_fluid*)handle(env, a.argv[0]), &m, &h, &sp, &rho));
  napi_value obj;
  napi_create_object(env, &obj);
  set(env, obj, "mass", number(env, m));
  set(env, obj, "smoothingRadius", number(env, h));
  set(env, obj, "particleSpacing", number(env, sp));
  set(env, obj, "restDensity", number(env, rho));
  return obj;
}
napi_value FluidDestroy(napi_env env, napi_callback_info info) {
  Args a = get_args(env, info);
  PF_CALL(pf_fluid_destroy((pf_fluid*)handle(env, a.argv[0])));
  return nullptr;
}

// ---------------------------------------------------------------- primitives
napi_value PrimitivesCreate(napi_env env, napi_callback_info info) {
  Args a = get_args(env, info);
  pf_primitives* s = nullptr;
  PF_CALL(pf_primitives_create((pf_particles*)handle(env, a.argv[0]), u32(env, a.argv[1]), &s));
  return make_handle(env, s);
}
// (set, kind, ...) kind: 'plane' n p muS muK | 'sphere' c r invert muS muK |
// 'box' c half rot invert muS muK | 'capsule' a b r muS muK
napi_value PrimitivesAdd(napi_env env, napi_callback_info info) {
  size_t argc = 8;
  napi_value v[8];
  napi_get_cb_info(env, info, &argc, v, nullptr, nullptr);
  for (size_t i = argc; i < 8; ++i) napi_get_undefined(env, &v[i]);
  pf_primitives* s = (pf_primitives*)handle(env, v[0]);
  char kind[16] = {0};
  size_t len = 0;
  napi_get_value_string_utf8(env, v[1], kind, sizeof kind, &len);
  uint32_t slot = 0;
  std::string k(kind);
  if (k == "plane") {
    auto n = doubles(env, v[2]), p = doubles(env, v[3]);
    PF_CALL(pf_primitives_add_plane(s, n.data(), p.data(), num(env, v[4], -1), num(env, v[5], -1), &slot));
  } else if (k == "sphere") {
    auto c = doubles(env, v[2]);
    bool inv = false;
    if (!is_nullish(env, v[4])) napi_get_value_bool(env, v[4], &inv);
    PF_CALL(pf_primitives_add_sphere(s, c.data(), num(env, v[3]), inv, num(env, v[5], -1), num(env, v[6], -1), &slot));
  } else if (k == "box") {
    auto c = doubles(env, v[2]), h = doubles(env, v[3]), q = doubles(env, v[4]);
    bool inv = false;
    if (!is_nullish(env, v[5])) napi_get_value_bool(env, v[5], &inv);
    PF_CALL(pf_primitives_add_box(s, c.data(), h.data(), q.size() == 4 ? q.data() : nullptr, inv, num(env, v[6], -1),
                                  num(env, v[7], -1), &slot));
  } else if (k == "capsule") {
    auto p0 = doubles(env, v[2]), p1 = doubles(env, v[3]);
    PF_CALL(pf_primitives_add_capsule(s, p0.data(), p1.data(), num(env, v[4]), num(env, v[5], -1), num(env, v[6], -1), &slot));
  } else {
    napi_throw_error(env, nullptr, "PrimitiveSet: unknown primitive kind");
    return nullptr;
  }
  return number(env, slot);
}
napi_value PrimitivesSetSphere(napi_env env, napi_callback_info info) {
  Args a = get_args(env, info);
  auto c = doubles(env, a.argv[2]), vel = doubles(env, a.argv[4]);
  PF_CALL(pf_primitives_set_sphere((pf_primitives*)handle(env, a.argv[0]), u32(env, a.argv[1]), c.data(),
                                   num(env, a.argv[3]), vel.size() == 3 ? vel.data() : nullptr));
  return nullptr;
}
napi_value PrimitivesSetPose(napi_env env, napi_callback_info info) {
  Args a = get_args(env, info);
  auto p = doubles(env, a.argv[2]), q = doubles(env, a.argv[3]);
  PF_CALL(pf_primitives_set_pose((pf_primitives*)handle(env, a.argv[0]), u32(env, a.argv[1]), p.data(),
                                 q.size() == 4 ? q.data() : nullptr, num(env, a.argv[4])));
  return nullptr;
}
napi_value PrimitivesCount(napi_env env, napi_callback_info info) {
  Args a = get_args(env, info);
  uint32_t n = 0;
  PF_CALL(pf_primitives_count((pf_primitives*)handle(env, a.argv[0]), &n));
  return number(env, n);
}
napi_value PrimitivesDestroy(napi_env env, napi_callback_info info) {
  Args a = get_args(env, info);
  PF_CALL(pf_primitives_destroy((pf_primitives*)handle(env, a.argv[0])));
  return nullptr;
}

// ---------------------------------------------------------------- sdf collider
// sdfCreate(particles, data: Float32Array, resolution[3], origin[3], voxelSize[3],
//           {muS, muK, position, rotation, scale, thickness})
napi_value SdfCreate(napi_env env, napi_callback_info info) {
  Args a = get_args(env, info);
  size_t nd = 0;
  const float* data = typed<float>(env, a.argv[1], &nd);
  if (pending(env)) return nullptr;
  auto res = doubles(env, a.argv[2]), origin = doubles(env, a.argv[3]), voxel = doubles(env, a.argv[4]);
  if (res.size() != 3 || origin.size() != 3 || voxel.size() != 3) {
    napi_throw_error(env, nullptr, "SDFCollider: resolution, origin and voxelSize take three numbers");
    return nullptr;
  }
  if (!(res[0] >= 0 && res[1] >= 0 && res[2] >= 0) || (double)nd != res[0] * res[1] * res[2]) {
    return throw_range(env, "SDFCollider: data holds %zu values, resolution %gx%gx%g needs %g", nd, res[0], res[1],
                       res[2], res[0] * res[1] * res[2]);
  }
  const uint32_t r3[3] = {(uint32_t)res[0], (uint32_t)res[1], (uint32_t)res[2]};
  if ((size_t)r3[0] * r3[1] * r3[2] != nd) {
    std::string msg = "SDFCollider: " + std::to_string(nd) + " values do not match resolution " +
                      std::to_string(r3[0]) + "×" + std::to_string(r3[1]) + "×" + std::to_string(r3[2]);
    napi_throw_error(env, nullptr, msg.c_str());
    return nullptr;
  }
  pf_sdf_options o{};
  napi_value opts = a.argv[5];
  o.mu_s = -1;
  o.mu_k = -1;
  std::vector<double> q;
  if (!is_nullish(env, opts)) {
    napi_value v = prop(env, opts, "muS");
    if (!is_nullish(env, v)) o.mu_s = num(env, v);
    v = prop(env, opts, "muK");
    if (!is_nullish(env, v)) o.mu_k = num(env, v);
    auto pos = doubles(env, prop(env, opts, "position"));
    if (pos.size() == 3) for (int k = 0; k < 3; ++k) o.position[k] = pos[k];
    q = doubles(env, prop(env, opts, "rotation"));
    if (q.size() == 4) o.rotation = q.data();
    o.scale = num(env, prop(env, opts, "scale"), 0);
    o.thickness = num(env, prop(env, opts, "thickness"), 0);
  }
  pf_sdf_collider* c = nullptr;
  PF_CALL(pf_sdf_collider_create((pf_particles*)handle(env, a.argv[0]), data, r3, origin.data(), voxel.data(), &o, &c));
  return make_handle(env, c);
}
napi_value SdfSet(napi_env env, napi_callback_info info) {
  Args a = get_args(env, info);
  auto* c = (pf_sdf_collider*)handle(env, a.argv[0]);
  char buf[16] = {0};
  size_t len = 0;
  napi_get_value_string_utf8(env, a.argv[1], buf, sizeof buf, &len);
  const std::string what(buf);
  if (what == "position") {
    auto p = doubles(env, a.argv[2]);
    if (p.size() != 3) { napi_throw_error(env, nullptr, "SDFCollider.setPosition: three numbers"); return nullptr; }
    PF_CALL(pf_sdf_collider_set_position(c, p.data()));
  } else if (what == "rotation") {
    auto q = doubles(env, a.argv[2]);
    if (q.size() != 4) { napi_throw_error(env, nullptr, "SDFCollider.setRotation: four numbers"); return nullptr; }
    PF_CALL(pf_sdf_collider_set_rotation(c, q.data()));
  } else if (what == "scale") {
    PF_CALL(pf_sdf_collider_set_scale(c, num(env, a.argv[2])));
  } else if (what == "friction") {
    PF_CALL(pf_sdf_collider_set_friction(c, num(env, a.argv[2]), num(env, a.argv[3])));
  } else if (what == "thickness") {
    PF_CALL(pf_sdf_collider_set_thickness(c, num(env, a.argv[2])));
  }
  return nullptr;
}
napi_value SdfDestroy(napi_env env, napi_callback_info info) {
  Args a = get_args(env, info);
  PF_CALL(pf_sdf_collider_destroy((pf_sdf_collider*)handle(env, a.argv[0])));
  return nullptr;
}

// ---------------------------------------------------------------- softbody
napi_value SoftbodyCreate(napi_env env, napi_callback_info info) {
  Args a = get_args(env, info);
  uint32_t nb = 0;
  napi_get_array_length(env, a.argv[1], &nb);
  std::vector<pf_softbody_def> defs(nb);
  for (uint32_t b = 0; b < nb; ++b) {
    napi_value o;
    napi_get_element(env, a.argv[1], b, &o);
    defs[b].start = u32(env, prop(env, o, "start"));
    defs[b].count = u32(env, prop(env, o, "count"));
    napi_value sc = prop(env, o, "surfaceCount");
    defs[b].surface_count = is_nullish(env, sc) ? 0xffffffffu : u32(env, sc);
    defs[b].compliance = num(env, prop(env, o, "compliance"), 0);
    size_t nr = 0, ne = 0;
    defs[b].rest_positions = typed<float>(env, prop(env, o, "restPositions"), &nr);
    defs[b].rest_count = (uint32_t)nr;
    if (pending(env)) return nullptr;
    defs[b].edges = typed<uint32_t>(env, prop(env, o, "edges"), &ne);
    if (pending(env)) return nullptr;
    if (ne % 2 != 0) return throw_range(env, "SoftbodySystem: body %u has an odd number of edge indices", b);
    defs[b].edge_count = (uint32_t)(ne / 2);
  }
  const uint32_t mode = u32(env, a.argv[2], PF_SHAPE_MATCHING_GLOBAL);
  pf_softbody* sb = nullptr;
  PF_CALL(pf_softbody_create((pf_particles*)handle(env, a.argv[0]), defs.data(), nb, mode, &sb));
  return make_handle(env, sb);
}
napi_value SoftbodySetSelfCollision(napi_env env, napi_callback_info info) {
  Args a = get_args(env, info);
  bool b = false;
  napi_get_value_bool(env, a.argv[1], &b);
  PF_CALL(pf_softbody_set_self_collision((pf_softbody*)handle(env, a.argv[0]), b ? 1 : 0));
  return nullptr;
}
napi_value SoftbodySetCompliance(napi_env env, napi_callback_info info) {
  Args a = get_args(env, info);
  PF_CALL(pf_softbody_set_compliance((pf_softbody*)handle(env, a.argv[0]), u32(env, a.argv[1]), num(env, a.argv[2])));
  return nullptr;
}
napi_value SoftbodyReadbackFrames(napi_env env, napi_callback_info info) {
  Args a = get_args(env, info);
  uint32_t nb = u32(env, a.argv[1]);
  float *c, *r;
  napi_value obj;
  napi_create_object(env, &obj);
  set(env, obj, "bodyCenters", f32_array(env, 4 * (size_t)nb, &c));
  set(env, obj, "bodyRotations", f32_array(env, 12 * (size_t)nb, &r));
  PF_CALL(pf_softbody_readback_frames((pf_softbody*)handle(env, a.argv[0]), c, r));
  return obj;
}
napi_value SoftbodyDestroy(napi_env env, napi_callback_info info) {
  Args a = get_args(env, info);
  PF_CALL(pf_softbody_destroy((pf_softbody*)handle(env, a.argv[0])));
  return nullptr;
}

// ---------------------------------------------------------------- loop
// (ps, {substeps, iterations, gravity, hashOrigin}, [{kind, handle}], [colliders])
napi_value LoopCreate(napi_env env, napi_callback_info info) {
  Args a = get_args(env, info);
  pf_loop_options o{};
  napi_value opts = a.argv[1];
  o.substeps = u32(env, prop(env, opts, "substeps"), 0);
  napi_value it = prop(env, opts, "iterations");
  o.iterations = is_nullish(env, it) ? 0xffffffffu : u32(env, it);
  auto g = doubles(env, prop(env, opts, "gravity"));
  if (g.size() == 3) {
    o.gravity[0] = g[0]; o.gravity[1] = g[1]; o.gravity[2] = g[2];
  } else {
    o.gravity[1] = -9.81;
  }
  auto ho = doubles(env, prop(env, opts, "hashOrigin"));
  if (ho.size() == 3) {
    o.has_hash_origin = 1;
    o.hash_origin[0] = ho[0]; o.hash_origin[1] = ho[1]; o.hash_origin[2] = ho[2];
  }
  // contact: {muS, muK, maxContacts} (defaults applied natively); quiet:
  // flattened [start, count] ranges of materials' noSelfContacts
  napi_value ct = prop(env, opts, "contact");
  size_t nq = 0;
  const uint32_t* quiet = typed<uint32_t>(env, prop(env, opts, "quiet"), &nq);
  if (pending(env)) return nullptr;
  if (!is_nullish(env, ct)) {
    o.contact = 1;
    o.contact_mu_s = num(env, prop(env, ct, "muS"), -1.0);
    o.contact_mu_k = num(env, prop(env, ct, "muK"), -1.0);
    o.max_contacts = u32(env, prop(env, ct, "maxContacts"), 0);
    o.nquiet = (uint32_t)(nq / 2);
    o.quiet = quiet;
  }
  uint32_t nm = 0, nc = 0;
  if (!is_nullish(env, a.argv[2])) napi_get_array_length(env, a.argv[2], &nm);
  if (!is_nullish(env, a.argv[3])) napi_get_array_length(env, a.argv[3], &nc);
  std::vector<pf_material> mats(nm);
  for (uint32_t i = 0; i < nm; ++i) {
    napi_value m;
    napi_get_element(env, a.argv[2], i, &m);
    mats[i].kind = u32(env, prop(env, m, "kind"));
    mats[i].handle = handle(env, prop(env, m, "handle"));
  }
  std::vector<pf_collider> cols(nc);
  for (uint32_t i = 0; i < nc; ++i) {
    napi_value c;
    napi_get_element(env, a.argv[3], i, &c);
    cols[i].kind = u32(env, prop(env, c, "kind"));
    cols[i].handle = handle(env, prop(env, c, "handle"));
  }
  pf_loop* L = nullptr;
  PF_CALL(pf_loop_create((pf_particles*)handle(env, a.argv[0]), &o, mats.data(), nm, cols.data(), nc, &L));
  return make_handle(env, L);
}
napi_value LoopStep(napi_env env, napi_callback_info info) {
  Args a = get_args(env, info);
  PF_CALL(pf_loop_step((pf_loop*)handle(env, a.argv[0]), num(env, a.argv[1])));
  return nullptr;
}

struct StepWork {
  pf_fence* fence;
  napi_deferred deferred;
  napi_async_work work;
  int rc;
  std::string error;
};

void step_execute(napi_env, void* data) {
  // Waits on the step's fence only: the loop and its stream stay the main
  // thread's (it may step or dispose meanwhile).
  StepWork* w = static_cast<StepWork*>(data);
  w->rc = pf_fence_wait(w->fence);
  if (w->rc != PF_OK) w->error = pf_last_error();
}

void step_complete(napi_env env, napi_status, void* data) {
  StepWork* w = static_cast<StepWork*>(data);
  if (w->rc == PF_OK) {
    napi_value u;
    napi_get_undefined(env, &u);
    napi_resolve_deferred(env, w->deferred, u);
  } else {
    napi_value msg, err;
    napi_create_string_utf8(env, w->error.c_str(), NAPI_AUTO_LENGTH, &msg);
    napi_create_error(env, nullptr, msg, &err);
    napi_reject_deferred(env, w->deferred, err);
  }
  napi_delete_async_work(env, w->work);
  pf_fence_destroy(w->fence);
  delete w;
}

// Enqueue a step and return a Promise resolved when the GPU has finished it.
napi_value LoopStepAsync(napi_env env, napi_callback_info info) {
  Args a = get_args(env, info);
  pf_loop* L = (pf_loop*)handle(env, a.argv[0]);
  PF_CALL(pf_loop_step(L, num(env, a.argv[1])));
  pf_fence* fence = nullptr;
  PF_CALL(pf_loop_fence(L, &fence));
  auto* w = new StepWork{fence, nullptr, nullptr, PF_OK, {}};
  napi_value promise, name;
  CHECK_NAPI(napi_create_promise(env, &w->deferred, &promise));
  napi_create_string_utf8(env, "pf_loop_step", NAPI_AUTO_LENGTH, &name);
  CHECK_NAPI(napi_create_async_work(env, nullptr, name, step_execute, step_complete, w, &w->work));
  CHECK_NAPI(napi_queue_async_work(env, w->work));
  return promise;
}
napi_value LoopSynchronize(napi_env env, napi_callback_info info) {
  Args a = get_args(env, info);
  PF_CALL(pf_loop_synchronize((pf_loop*)handle(env, a.argv[0])));
  return nullptr;
}
napi_value LoopReadbackOverflow(napi_env env, napi_callback_info info) {
  Args a = get_args(env, info);
  pf_overflow o{};
  PF_CALL(pf_loop_readback_overflow((pf_loop*)handle(env, a.argv[0]), &o));
  napi_value obj;
  napi_create_object(env, &obj);
  set(env, obj, "positions", boolean(env, o.positions));
  set(env, obj, "velocities", boolean(env, o.velocities));
  set(env, obj, "grid", boolean(env, o.grid));
  set(env, obj, "contacts", boolean(env, o.contacts));
  return obj;
}
// ---------------------------------------------------------------- formats
napi_value num_array(napi_env env, const double* v, uint32_t n) {
  napi_value arr;
  napi_create_array_with_length(env, n, &arr);
  for (uint32_t i = 0; i < n; ++i) napi_set_element(env, arr, i, number(env, v[i]));
  return arr;
}
// (bytes: Uint8Array) -> SDFData {data, resolution, origin, voxelSize}
napi_value SdfDecode(napi_env env, napi_callback_info info) {
  Args a = get_args(env, info);
  size_t n = 0;
  const uint8_t* b = typed<uint8_t>(env, a.argv[0], &n);
  if (pending(env)) return nullptr;
  uint32_t res[3] = {0, 0, 0};
  double o[3], v[3];
  size_t count = 0;
  PF_CALL(pf_sdf_decode(b, n, res, o, v, nullptr, 0, &count));
  float* d = nullptr;
  napi_value data = f32_array(env, count, &d);
  PF_CALL(pf_sdf_decode(b, n, res, o, v, d, count, &count));
  const double rd[3] = {(double)res[0], (double)res[1], (double)res[2]};
  napi_value obj;
  napi_create_object(env, &obj);
  set(env, obj, "data", data);
  set(env, obj, "resolution", num_array(env, rd, 3));
  set(env, obj, "origin", num_array(env, o, 3));
  set(env, obj, "voxelSize", num_array(env, v, 3));
  return obj;
}
// (data: Float32Array, resolution, origin, voxelSize) -> Uint8Array
napi_value SdfEncode(napi_env env, napi_callback_info info) {
  Args a = get_args(env, info);
  size_t count = 0;
  const float* d = typed<float>(env, a.argv[0], &count);
  if (pending(env)) return nullptr;
  auto r = doubles(env, a.argv[1]), o = doubles(env, a.argv[2]), v = doubles(env, a.argv[3]);
  if (r.size() != 3 || o.size() != 3 || v.size() != 3) {
    napi_throw_type_error(env, nullptr, "encodeSdfBinary: resolution, origin and voxelSize need three numbers");
    return nullptr;
  }
  const uint32_t res[3] = {(uint32_t)r[0], (uint32_t)r[1], (uint32_t)r[2]};
  size_t size = 0;
  PF_CALL(pf_sdf_encode(d, count, res, o.data(), v.data(), nullptr, 0, &size));
  napi_value ab, arr;
  void* p = nullptr;
  napi_create_arraybuffer(env, size, &p, &ab);
  napi_create_typedarray(env, napi_uint8_array, size, ab, 0, &arr);
  PF_CALL(pf_sdf_encode(d, count, res, o.data(), v.data(), p, size, &size));
  return arr;
}

// (loop, cap) -> {count, pairs: Uint32Array of the stored (i, j) pairs}
napi_value LoopReadbackContacts(napi_env env, napi_callback_info info) {
  Args a = get_args(env, info);
  const uint32_t cap = u32(env, a.argv[1], 0);
  std::vector<uint32_t> pairs(2 * (size_t)cap);
  uint32_t count = 0, stored = 0;
  PF_CALL(pf_loop_readback_contacts((pf_loop*)handle(env, a.argv[0]), &count, cap ? pairs.data() : nullptr, cap,
                                    &stored));
  uint32_t* out;
  napi_value arr = u32_array(env, 2 * (size_t)stored, &out);
  std::copy(pairs.begin(), pairs.begin() + 2 * (size_t)stored, out);
  napi_value obj;
  napi_create_object(env, &obj);
  set(env, obj, "count", number(env, count));
  set(env, obj, "pairs", arr);
  return obj;
}
napi_value LoopSetSubsteps(napi_env env, napi_callback_info info) {
  Args a = get_args(env, info);
  PF_CALL(pf_loop_set_substeps((pf_loop*)handle(env, a.argv[0]), u32(env, a.argv[1])));
  return nullptr;
}
napi_value LoopSetGravity(napi_env env, napi_callback_info info) {
  Args a = get_args(env, info);
  auto g = doubles(env, a.argv[1]);
  if (g.size() != 3) {
    napi_throw_type_error(env, nullptr, "SimLoop: gravity needs three numbers");
    return nullptr;
  }
  PF_CALL(pf_loop_set_gravity((pf_loop*)handle(env, a.argv[0]), g.data()));
  return nullptr;
}
napi_value LoopSetGraph(napi_env env, napi_callback_info info) {
  Args a = get_args(env, info);
  bool on = false;
  napi_get_value_bool(env, a.argv[1], &on);
  PF_CALL(pf_loop_set_graph((pf_loop*)handle(env, a.argv[0]), on ? 1 : 0));
  return nullptr;
}
napi_value LoopInfo(napi_env env, napi_callback_info info) {
  Args a = get_args(env, info);
  uint32_t s, it, H, k;
  double cell;
  PF_CALL(pf_loop_info((pf_loop*)handle(env, a.argv[0]), &s, &it, &cell, &H, &k));
  napi_value obj;
  napi_create_object(env, &obj);
  set(env, obj, "substeps", number(env, s));
  set(env, obj, "iterations", number(env, it));
  set(env, obj, "cellSize", number(env, cell));
  set(env, obj, "hashTableSize", number(env, H));
  set(env, obj, "kernelsPerStep", number(env, k));
  return obj;
}
napi_value LoopGrid(napi_env env, napi_callback_info info) {
  Args a = get_args(env, info);
  pf_grid* g = nullptr;
  PF_CALL(pf_loop_grid((pf_loop*)handle(env, a.argv[0]), &g));
  return make_handle(env, g);
}
napi_value LoopDestroy(napi_env env, napi_callback_info info) {
  Args a = get_args(env, info);
  PF_CALL(pf_loop_destroy((pf_loop*)handle(env, a.argv[0])));
  return nullptr;
}

// ---------------------------------------------------------------- viscosity solver
// (fluid, viscosity, iterations)
napi_value ViscosityCreate(napi_env env, napi_callback_info info) {
  Args a = get_args(env, info);
  pf_viscosity* v = nullptr;
  PF_CALL(pf_viscosity_create((pf_fluid*)handle(env, a.argv[0]), num(env, a.argv[1]), u32(env, a.argv[2], 12), &v));
  return make_handle(env, v);
}
napi_value ViscositySet(napi_env env, napi_callback_info info) {
  Args a = get_args(env, info);
  PF_CALL(pf_viscosity_set((pf_viscosity*)handle(env, a.argv[0]), num(env, a.argv[1])));
  return nullptr;
}
napi_value ViscosityDestroy(napi_env env, napi_callback_info info) {
  Args a = get_args(env, info);
  PF_CALL(pf_viscosity_destroy((pf_viscosity*)handle(env, a.argv[0])));
  return nullptr;
}

// ---------------------------------------------------------------- surface field
// (fluid, {min, max, voxelBudget, motionStretch, solids: {start, count}, wet,
//  radius, meniscusHeight, meniscusWidth}) -> {handle, resolution, origin, voxel}
napi_value SurfaceFieldCreate(napi_env env, napi_callback_info info) {
  Args a = get_args(env, info);
  napi_value o = a.argv[1];
  pf_surface_field_options opt{};
  auto mn = doubles(env, prop(env, o, "min")), mx = doubles(env, prop(env, o, "max"));
  if (mn.size() != 3 || mx.size() != 3) {
    napi_throw_type_error(env, nullptr, "FluidSurfaceRenderer: bounds need min and max of three numbers");
    return nullptr;
  }
  for (int k = 0; k < 3; ++k) {
    opt.min[k] = mn[k];
    opt.max[k] = mx[k];
  }
  opt.voxel_budget = num(env, prop(env, o, "voxelBudget"), 0);
  opt.motion_stretch = num(env, prop(env, o, "motionStretch"), 0);
  napi_value sol = prop(env, o, "solids");
  if (!is_nullish(env, sol)) {
    opt.solids_start = u32(env, prop(env, sol, "start"));
    opt.solids_count = u32(env, prop(env, sol, "count"));
  }
  bool wet = false;
  napi_value w = prop(env, o, "wet");
  if (!is_nullish(env, w)) napi_get_value_bool(env, w, &wet);
  opt.wet = wet ? 1 : 0;
  opt.radius = num(env, prop(env, o, "radius"), -1);
  opt.meniscus_height = num(env, prop(env, o, "meniscusHeight"), -1);
  opt.meniscus_width = num(env, prop(env, o, "meniscusWidth"), -1);
  pf_surface_field* f = nullptr;
  PF_CALL(pf_surface_field_create((pf_fluid*)handle(env, a.argv[0]), &opt, &f));
  uint32_t res[3];
  double org[3], voxel = 0;
  PF_CALL(pf_surface_field_info(f, res, org, &voxel));
  const double rd[3] = {(double)res[0], (double)res[1], (double)res[2]};
  napi_value obj;
  napi_create_object(env, &obj);
  set(env, obj, "handle", make_handle(env, f));
  set(env, obj, "resolution", num_array(env, rd, 3));
  set(env, obj, "origin", num_array(env, org, 3));
  set(env, obj, "voxel", number(env, voxel));
  return obj;
}
napi_value SurfaceFieldUpdate(napi_env env, napi_callback_info info) {
  Args a = get_args(env, info);
  PF_CALL(pf_surface_field_update((pf_surface_field*)handle(env, a.argv[0])));
  return nullptr;
}
// (handle, voxelCount) -> Float32Array of 4 floats per voxel
napi_value SurfaceFieldReadback(napi_env env, napi_callback_info info) {
  Args a = get_args(env, info);
  float* d = nullptr;
  napi_value arr = f32_array(env, 4 * (size_t)num(env, a.argv[1]), &d);
  PF_CALL(pf_surface_field_readback((pf_surface_field*)handle(env, a.argv[0]), d));
  return arr;
}
napi_value SurfaceFieldDestroy(napi_env env, napi_callback_info info) {
  Args a = get_args(env, info);
  PF_CALL(pf_surface_field_destroy((pf_surface_field*)handle(env, a.argv[0])));
  return nullptr;
}

// ---------------------------------------------------------------- autotuner
// (handle, dt, [candidates], steps) -> {best, msPerStep: {capacity: ms}}
template <typename Fn>
napi_value autotune(napi_env env, napi_callback_info info, Fn fn) {
  Args a = get_args(env, info);
  auto c = doubles(env, a.argv[2]);
  std::vector<uint32_t> cand(c.begin(), c.end());
  std::vector<float> ms(cand.size(), 0.f);
  uint32_t best = 0;
  PF_CALL(fn(handle(env, a.argv[0]), num(env, a.argv[1]), cand.data(), (uint32_t)cand.size(),
             u32(env, a.argv[3], 2), &best, ms.data()));
  napi_value obj, per;
  napi_create_object(env, &obj);
  napi_create_object(env, &per);
  for (size_t i = 0; i < cand.size(); ++i) set(env, per, std::to_string(cand[i]).c_str(), number(env, ms[i]));
  set(env, obj, "best", number(env, best));
  set(env, obj, "msPerStep", per);
  return obj;
}
napi_value LoopAutotune(napi_env env, napi_callback_info info) {
  return autotune(env, info, [](void* h, double dt, const uint32_t* c, uint32_t n, uint32_t st, uint32_t* b,
                                float* ms) { return pf_loop_autotune((pf_loop*)h, dt, c, n, st, b, ms); });
}

// ---------------------------------------------------------------- batched form
// Resolves a Promise from a worker thread once the GPU has passed a fence
// recorded on the context's stream (the worker never touches the stream).
struct SyncWork {
  pf_fence* fence;
  napi_deferred deferred;
  napi_async_work work;
  int rc;
  std::string error;
};

napi_value promise_after_sync(napi_env env, pf_context* ctx) {
  pf_fence* fence = nullptr;
  PF_CALL(pf_context_fence(ctx, &fence));
  auto* w = new SyncWork{fence, nullptr, nullptr, PF_OK, {}};
  napi_value promise, name;
  CHECK_NAPI(napi_create_promise(env, &w->deferred, &promise));
  napi_create_string_utf8(env, "pf_context_synchronize", NAPI_AUTO_LENGTH, &name);
  CHECK_NAPI(napi_create_async_work(
      env, nullptr, name,
      [](napi_env, void* d) {
        auto* w = static_cast<SyncWork*>(d);
        w->rc = pf_fence_wait(w->fence);
        if (w->rc != PF_OK) w->error = pf_last_error();
      },
      [](napi_env env, napi_status, void* d) {
        auto* w = static_cast<SyncWork*>(d);
        if (w->rc == PF_OK) {
          napi_value u;
          napi_get_undefined(env, &u);
          napi_resolve_deferred(env, w->deferred, u);
        } else {
          napi_value msg, err;
          napi_create_string_utf8(env, w->error.c_str(), NAPI_AUTO_LENGTH, &msg);
          napi_create_error(env, nullptr, msg, &err);
          napi_reject_deferred(env, w->deferred, err);
        }
        napi_delete_async_work(env, w->work);
        pf_fence_destroy(w->fence);
        delete w;
      },
      w, &w->work));
  CHECK_NAPI(napi_queue_async_work(env, w->work));
  return promise;
}

// (ctx, [loops])
napi_value BatchCreate(napi_env env, napi_callback_info info) {
  Args a = get_args(env, info);
  uint32_t n = 0;
  napi_get_array_length(env, a.argv[1], &n);
  std::vector<pf_loop*> loops(n);
  for (uint32_t i = 0; i < n; ++i) {
    napi_value e;
    napi_get_element(env, a.argv[1], i, &e);
    loops[i] = (pf_loop*)handle(env, e);
  }
  pf_batch* b = nullptr;
  PF_CALL(pf_batch_create((pf_context*)handle(env, a.argv[0]), loops.data(), n, &b));
  return make_handle(env, b);
}
// (ctx, batch, dt) -> Promise resolved when the step has run
napi_value BatchStepAsync(napi_env env, napi_callback_info info) {
  Args a = get_args(env, info);
  PF_CALL(pf_batch_step((pf_batch*)handle(env, a.argv[1]), num(env, a.argv[2])));
  return promise_after_sync(env, (pf_context*)handle(env, a.argv[0]));
}
napi_value BatchSetGraph(napi_env env, napi_callback_info info) {
  Args a = get_args(env, info);
  bool on = false;
  napi_get_value_bool(env, a.argv[1], &on);
  PF_CALL(pf_batch_set_graph((pf_batch*)handle(env, a.argv[0]), on ? 1 : 0));
  return nullptr;
}
napi_value BatchAutotune(napi_env env, napi_callback_info info) {
  return autotune(env, info, [](void* h, double dt, const uint32_t* c, uint32_t n, uint32_t st, uint32_t* b,
                                float* ms) { return pf_batch_autotune((pf_batch*)h, dt, c, n, st, b, ms); });
}
napi_value BatchDestroy(napi_env env, napi_callback_info info) {
  Args a = get_args(env, info);
  PF_CALL(pf_batch_destroy((pf_batch*)handle(env, a.argv[0])));
  return nullptr;
}

// ---------------------------------------------------------------- served form
napi_value request_info(napi_env env, const pf_request_info& r) {
  static const char* status[] = {"pending", "launched", "done", "failed"};
  napi_value obj, st;
  napi_create_object(env, &obj);
  set(env, obj, "scene", number(env, r.scene));
  napi_create_string_utf8(env, status[r.status & 3u], NAPI_AUTO_LENGTH, &st);
  set(env, obj, "status", st);
  set(env, obj, "launchScenes", number(env, r.launch_scenes));
  set(env, obj, "tSubmit", number(env, r.t_submit));
  set(env, obj, "tLaunch", number(env, r.t_launch));
  set(env, obj, "tDone", number(env, r.t_done));
  return obj;
}

// (ctx, {hbmBudget, maxBatchScenes, maxBatchParticles, slots, hostBudget})
napi_value ServerCreate(napi_env env, napi_callback_info info) {
  Args a = get_args(env, info);
  pf_server_options o{};
  napi_value opts = a.argv[1];
  if (!is_nullish(env, opts)) {
    o.hbm_budget_bytes = (uint64_t)num(env, prop(env, opts, "hbmBudget"), 0);
    o.max_batch_scenes = u32(env, prop(env, opts, "maxBatchScenes"), 0);
    o.max_batch_particles = (uint64_t)num(env, prop(env, opts, "maxBatchParticles"), 0);
    o.slots = u32(env, prop(env, opts, "slots"), 0);
    o.host_budget_bytes = (uint64_t)num(env, prop(env, opts, "hostBudget"), 0);
  }
  pf_server* s = nullptr;
  PF_CALL(pf_server_create((pf_context*)handle(env, a.argv[0]), &o, &s));
  return make_handle(env, s);
}
napi_value ServerAddScene(napi_env env, napi_callback_info info) {
  Args a = get_args(env, info);
  uint32_t id = 0;
  PF_CALL(pf_server_add_scene((pf_server*)handle(env, a.argv[0]), (pf_loop*)handle(env, a.argv[1]), &id));
  return number(env, id);
}
napi_value ServerRemoveScene(napi_env env, napi_callback_info info) {
  Args a = get_args(env, info);
  PF_CALL(pf_server_remove_scene((pf_server*)handle(env, a.argv[0]), u32(env, a.argv[1])));
  return nullptr;
}
napi_value ServerSubmit(napi_env env, napi_callback_info info) {
  Args a = get_args(env, info);
  uint64_t t = 0;
  PF_CALL(pf_server_submit((pf_server*)handle(env, a.argv[0]), u32(env, a.argv[1]), num(env, a.argv[2]), &t));
  return number(env, (double)t);
}

// Blocks in a worker thread on pf_server_wait (ticket) or pf_server_drain
// (ticket < 0) and settles the Promise on the JS thread.
struct WaitWork {
  pf_server* srv;
  double ticket;
  napi_deferred deferred;
  napi_async_work work;
  int rc;
  pf_request_info info;
  std::string error;
};

napi_value server_wait_promise(napi_env env, pf_server* srv, double ticket) {
  auto* w = new WaitWork{srv, ticket, nullptr, nullptr, PF_OK, {}, {}};
  napi_value promise, name;
  CHECK_NAPI(napi_create_promise(env, &w->deferred, &promise));
  napi_create_string_utf8(env, "pf_server_wait", NAPI_AUTO_LENGTH, &name);
  CHECK_NAPI(napi_create_async_work(
      env, nullptr, name,
      [](napi_env, void* d) {
        auto* w = static_cast<WaitWork*>(d);
        w->rc = w->ticket < 0 ? pf_server_drain(w->srv) : pf_server_wait(w->srv, (uint64_t)w->ticket, &w->info);
        if (w->rc != PF_OK) w->error = pf_last_error();
      },
      [](napi_env env, napi_status, void* d) {
        auto* w = static_cast<WaitWork*>(d);
        if (w->rc == PF_OK) {
          napi_value v;
          if (w->ticket < 0) napi_get_undefined(env, &v);
          else v = request_info(env, w->info);
          napi_resolve_deferred(env, w->deferred, v);
        } else {
          napi_value msg, err;
          napi_create_string_utf8(env, w->error.c_str(), NAPI_AUTO_LENGTH, &msg);
          napi_create_error(env, nullptr, msg, &err);
          napi_reject_deferred(env, w->deferred, err);
        }
        napi_delete_async_work(env, w->work);
        delete w;
      },
      w, &w->work));
  CHECK_NAPI(napi_queue_async_work(env, w->work));
  return promise;
}
napi_value ServerWait(napi_env env, napi_callback_info info) {
  Args a = get_args(env, info);
  return server_wait_promise(env, (pf_server*)handle(env, a.argv[0]), num(env, a.argv[1]));
}
napi_value ServerDrain(napi_env env, napi_callback_info info) {
  Args a = get_args(env, info);
  return server_wait_promise(env, (pf_server*)handle(env, a.argv[0]), -1.0);
}
napi_value ServerStats(napi_env env, napi_callback_info info) {
  Args a = get_args(env, info);
  pf_server_stats st{};
  PF_CALL(pf_server_stats_get((pf_server*)handle(env, a.argv[0]), &st));
  napi_value obj;
  napi_create_object(env, &obj);
#define PF_STAT(k, v) set(env, obj, k, number(env, (double)(v)))
  PF_STAT("submitted", st.submitted); PF_STAT("completed", st.completed); PF_STAT("failed", st.failed);
  PF_STAT("launches", st.launches); PF_STAT("sceneSteps", st.scene_steps);
  PF_STAT("particleSubsteps", st.particle_substeps); PF_STAT("evictions", st.evictions);
  PF_STAT("restores", st.restores); PF_STAT("residentBytes", st.resident_bytes);
  PF_STAT("peakResidentBytes", st.peak_resident_bytes); PF_STAT("hbmBudget", st.hbm_budget_bytes);
  PF_STAT("scenes", st.scenes); PF_STAT("residentScenes", st.resident_scenes);
  PF_STAT("maxLaunchScenes", st.max_launch_scenes);
  PF_STAT("hostBytes", st.host_bytes); PF_STAT("evictFailures", st.evict_failures);
#undef PF_STAT
  return obj;
}
napi_value ServerEvictScene(napi_env env, napi_callback_info info) {
  Args a = get_args(env, info);
  PF_CALL(pf_server_evict_scene((pf_server*)handle(env, a.argv[0]), u32(env, a.argv[1])));
  return nullptr;
}
// Launch shapes (no reference counterpart): 0 or a persistent workgroup cap.
napi_value LoopSetLaunchShape(napi_env env, napi_callback_info info) {
  Args a = get_args(env, info);
  PF_CALL(pf_loop_set_launch_shape((pf_loop*)handle(env, a.argv[0]), u32(env, a.argv[1])));
  return nullptr;
}
napi_value BatchSetLaunchShape(napi_env env, napi_callback_info info) {
  Args a = get_args(env, info);
  PF_CALL(pf_batch_set_launch_shape((pf_batch*)handle(env, a.argv[0]), u32(env, a.argv[1])));
  return nullptr;
}
napi_value ServerSetLaunchShape(napi_env env, napi_callback_info info) {
  Args a = get_args(env, info);
  PF_CALL(pf_server_set_launch_shape((pf_server*)handle(env, a.argv[0]), u32(env, a.argv[1])));
  return nullptr;
}
napi_value LoopAutotuneLaunch(napi_env env, napi_callback_info info) {
  return autotune(env, info, [](void* h, double dt, const uint32_t* c, uint32_t n, uint32_t st, uint32_t* b,
                                float* ms) { return pf_loop_autotune_launch((pf_loop*)h, dt, c, n, st, b, ms); });
}
napi_value BatchAutotuneLaunch(napi_env env, napi_callback_info info) {
  return autotune(env, info, [](void* h, double dt, const uint32_t* c, uint32_t n, uint32_t st, uint32_t* b,
                                float* ms) { return pf_batch_autotune_launch((pf_batch*)h, dt, c, n, st, b, ms); });
}
napi_value ServerDestroy(napi_env env, napi_callback_info info) {
  Args a = get_args(env, info);
  PF_CALL(pf_server_destroy((pf_server*)handle(env, a.argv[0])));
  return nullptr;
}

// ---------------------------------------------------------------- layout
napi_value LayoutFillBox(napi_env env, napi_callback_info info) {
  Args a = get_args(env, info);
  auto mn = doubles(env, a.argv[0]), mx = doubles(env, a.argv[1]);
  double sp = num(env, a.argv[2]);
  uint32_t n = 0;
  PF_CALL(pf_layout_fill_box(mn.data(), mx.data(), sp, nullptr, 0, &n));
  float* d;
  napi_value arr = f32_array(env, 3 * (size_t)n, &d);
  PF_CALL(pf_layout_fill_box(mn.data(), mx.data(), sp, d, n, &n));
  return arr;
}
napi_value LayoutFitBudget(napi_env env, napi_callback_info info) {
  Args a = get_args(env, info);
  auto c = doubles(env, a.argv[1]), boxes = doubles(env, a.argv[2]);
  double r = 0;
  uint32_t n = 0;
  PF_CALL(pf_layout_fit_budget(u32(env, a.argv[0]), c.size() == 6 ? c.data() : nullptr, boxes.data(),
                               (uint32_t)(boxes.size() / 6), &r, &n));
  napi_value obj;
  napi_create_object(env, &obj);
  set(env, obj, "radius", number(env, r));
  set(env, obj, "count", number(env, n));
  return obj;
}
napi_value AbiVersion(napi_env env, napi_callback_info) { return number(env, pf_abi_version()); }

napi_value Init(napi_env env, napi_value exports) {
  struct Fn {
    const char* name;
    napi_callback cb;
  } fns[] = {
      {"contextCreate", ContextCreate}, {"contextDestroy", ContextDestroy}, {"contextSynchronize", ContextSynchronize},
      {"particlesCreate", ParticlesCreate}, {"particlesDestroy", ParticlesDestroy},
      {"particlesUpload", ParticlesUpload}, {"particlesSetInvMass", ParticlesSetInvMass},
      {"particlesSetCollisionGroup", ParticlesSetCollisionGroup}, {"particlesReadback", ParticlesReadback},
      {"gridCreate", GridCreate}, {"gridRebuild", GridRebuild}, {"gridReadback", GridReadback},
      {"gridReadbackOverflow", GridReadbackOverflow}, {"gridInfo", GridInfo}, {"gridSetOrigin", GridSetOrigin},
      {"gridDestroy", GridDestroy},
      {"fluidCreate", FluidCreate}, {"fluidSet", FluidSet}, {"fluidGet", FluidGet},
      {"fluidAddBoundary", FluidAddBoundary}, {"fluidReadbackDensity", FluidReadbackDensity},
      {"fluidReadbackOverflow", FluidReadbackOverflow}, {"fluidInfo", FluidInfo}, {"fluidDestroy", FluidDestroy},
      {"primitivesCreate", PrimitivesCreate}, {"primitivesAdd", PrimitivesAdd},
      {"primitivesSetSphere", PrimitivesSetSphere}, {"primitivesSetPose", PrimitivesSetPose},
      {"primitivesCount", PrimitivesCount}, {"primitivesDestroy", PrimitivesDestroy},
      {"sdfCreate", SdfCreate}, {"sdfSet", SdfSet}, {"sdfDestroy", SdfDestroy},
      {"softbodyCreate", SoftbodyCreate}, {"softbodySetCompliance", SoftbodySetCompliance},
      {"softbodySetSelfCollision", SoftbodySetSelfCollision}, {"loopReadbackContacts", LoopReadbackContacts},
      {"sdfDecode", SdfDecode}, {"sdfEncode", SdfEncode},
      {"softbodyReadbackFrames", SoftbodyReadbackFrames}, {"softbodyDestroy", SoftbodyDestroy},
      {"loopCreate", LoopCreate}, {"loopStep", LoopStep}, {"loopStepAsync", LoopStepAsync},
      {"loopSynchronize", LoopSynchronize}, {"loopReadbackOverflow", LoopReadbackOverflow},
      {"loopSetSubsteps", LoopSetSubsteps}, {"loopSetGravity", LoopSetGravity}, {"loopSetGraph", LoopSetGraph},
      {"loopInfo", LoopInfo}, {"loopGrid", LoopGrid}, {"loopDestroy", LoopDestroy},
      {"loopAutotune", LoopAutotune}, {"particlesDeviceBytes", ParticlesDeviceBytes},
      {"viscosityCreate", ViscosityCreate}, {"viscositySet", ViscositySet}, {"viscosityDestroy", ViscosityDestroy},
      {"surfaceFieldCreate", SurfaceFieldCreate}, {"surfaceFieldUpdate", SurfaceFieldUpdate},
      {"surfaceFieldReadback", SurfaceFieldReadback}, {"surfaceFieldDestroy", SurfaceFieldDestroy},
      {"batchCreate", BatchCreate}, {"batchStepAsync", BatchStepAsync}, {"batchSetGraph", BatchSetGraph},
      {"batchAutotune", BatchAutotune}, {"batchDestroy", BatchDestroy},
      {"serverCreate", ServerCreate}, {"serverAddScene", ServerAddScene}, {"serverRemoveScene", ServerRemoveScene},
      {"serverSubmit", ServerSubmit}, {"serverWait", ServerWait}, {"serverDrain", ServerDrain},
      {"serverStats", ServerStats}, {"serverEvictScene", ServerEvictScene}, {"serverDestroy", ServerDestroy},
      {"loopSetLaunchShape", LoopSetLaunchShape}, {"batchSetLaunchShape", BatchSetLaunchShape},
      {"serverSetLaunchShape", ServerSetLaunchShape}, {"loopAutotuneLaunch", LoopAutotuneLaunch},
      {"batchAutotuneLaunch", BatchAutotuneLaunch},
      {"layoutFillBox", LayoutFillBox}, {"layoutFitBudget", LayoutFitBudget}, {"abiVersion", AbiVersion},
  };
  for (const Fn& f : fns) {
    napi_value fn;
    napi_create_function(env, f.name, NAPI_AUTO_LENGTH, f.cb, nullptr, &fn);
    napi_set_named_property(env, exports, f.name, fn);
  }
  return exports;
}

}  // namespace

NAPI_MODULE(NODE_GYP_MODULE_NAME, Init)
