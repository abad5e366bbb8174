// Host-side objects behind the C ABI.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "../../include/pf.h"
#include "pf_internal.hpp"
#include "pf_kernels.hpp"

namespace pf {

void set_error(const char* fmt, ...);

struct HipError {
  int code;
};

struct PfError {
  int code;
};

#define PF_HIP(expr)                                                               \
  do {                                                                             \
    hipError_t e__ = (expr);                                                       \
    if (e__ != hipSuccess) {                                                       \
      ::pf::set_error("HIP: %s failed: %s", #expr, hipGetErrorString(e__));       \
      throw ::pf::HipError{PF_EDEVICE};                                            \
    }                                                                              \
  } while (0)

// Throw an argument error with the reference's "<Owner>: <reason>" wording.
[[noreturn]] void fail(int code, const char* fmt, ...);

// A launch shape for the tile kernels: 0 (one workgroup per tile) or a
// persistent cap, a multiple of the 8 XCDs so each workgroup keeps its XCD's
// tiles, up to 2^20 workgroups.
inline void check_launch_shape(uint32_t workgroups, const char* owner) {
  if (workgroups % 8u != 0u || workgroups > (1u << 20))
    fail(PF_EINVAL, "%s: launch shape must be 0 (one workgroup per tile) or a multiple of 8 up to %u workgroups, "
         "got %u", owner, 1u << 20, workgroups);
}

// Device memory of one evictable scene. Each chunk is a range of reserved
// device addresses backed by one physical allocation; eviction copies the
// chunk's used bytes to pinned host memory and releases the physical memory,
// restore maps fresh physical memory at the same addresses and copies back.
// Pointers into the arena (descriptors, tile tables, captured graphs) stay
// valid across an evict/restore round trip. Buffers are bump-allocated and
// released together with the arena.
struct Arena {
  // A range of reserved device addresses backed, while resident, by one
  // physical allocation of the same size; buffers are bump-allocated in it.
  struct Segment {
    char* base = nullptr;
    size_t size = 0, used = 0;
    hipMemGenericAllocationHandle_t handle{};
    void* host = nullptr;  // pageable copy of the used bytes while evicted
  };
  int device = 0;
  size_t gran = 0;
  size_t hint = 0;  // expected bytes of the scene (sizes its first segment)
  std::vector<Segment> segs;
  bool resident = true;
  explicit Arena(int device, size_t hint_bytes = 0);
  ~Arena();
  void* alloc(size_t bytes);
  size_t bytes() const;       // physical bytes held while resident
  size_t used_bytes() const;  // bytes an eviction copies to host memory
  void evict();
  void restore();
};

// While alive, dalloc places buffers in `arena` (creation calls of an
// evictable scene); temporary buffers outside a scope use hipMalloc.
struct ArenaScope {
  Arena* prev;
  explicit ArenaScope(Arena* arena);
  ~ArenaScope();
};
Arena* current_arena();

void* dalloc_bytes(size_t bytes, bool temporary);

template <typename T>
T* dalloc(size_t n) {
  if (n == 0) n = 1;
  return static_cast<T*>(dalloc_bytes(n * sizeof(T), false));
}
// A short-lived buffer (uploads, readbacks): always hipMalloc, never an arena.
template <typename T>
T* dalloc_temp(size_t n) {
  if (n == 0) n = 1;
  return static_cast<T*>(dalloc_bytes(n * sizeof(T), true));
}
// Frees a hipMalloc'd buffer; arena buffers are released with their arena.
void dfree(void* p);

struct KernelStat {
  uint32_t launches = 0;
  double total_ms = 0.0;
};

struct Context;

// Brackets a launch with events when the context is profiling.
struct LaunchScope {
  Context* ctx;
  const char* name;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  LaunchScope(Context* c, const char* n);
  ~LaunchScope();
};
#define PF_LAUNCH(ctx, name) ::pf::LaunchScope pf_scope__(ctx, name)

struct Context {
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t t0 = nullptr, t1 = nullptr;
  bool profiling = false;
  bool capturing = false;
  std::vector<std::pair<std::string, std::pair<hipEvent_t, hipEvent_t>>> pending;
  std::vector<hipEvent_t> event_pool;
  std::map<std::string, KernelStat> stats;
  void* scan_ctl = nullptr;  // ScanCtlDev for launches on this stream
  hipEvent_t get_event();
  void collect();
};

// Per-scene device state shared by particles, grid, loop.
struct Particles {
  Context* ctx = nullptr;
  uint32_t n = 0;
  double radius = 0;
  bool disposed = false;
  bool stepped = false;  // state is cell-sorted (slot != id)
  SceneDev host{};       // host copy of the descriptor (buffer pairs unswapped)
  SceneDev* dev = nullptr;  // 4 device variants: bit 1 swaps state pairs, bit 0 swaps xs
  int variant = 0;       // variant whose [0] buffers hold the current state / prediction
  bool dirty = true;
  SceneDev desc(int v) const;
  SceneDev* cur() const { return dev + variant; }
  std::vector<float> cpu_inv_mass;   // host mirror of inverse masses (original order)
  std::vector<float> cpu_positions;  // host mirror of uploaded positions (xyz)
  std::vector<uint32_t> cpu_groups;
  struct pf_loop_s* loop = nullptr;  // attached loop, if any
  std::unique_ptr<Arena> arena;      // every device buffer of an evictable scene
  void* server = nullptr;            // the pf_server stepping this scene, if any
  void upload_desc();
};

struct Grid;
struct Fluid;
struct Primitives;
struct Softbody;
struct Loop;

}  // namespace pf

struct pf_context_s : pf::Context {};
struct pf_fence_s {
  hipEvent_t ev = nullptr;
};
struct pf_particles_s : pf::Particles {};

struct pf_grid_s {
  pf::Context* ctx = nullptr;
  pf_particles_s* ps = nullptr;  // null for standalone grids
  bool owned_by_loop = false;
  uint32_t n = 0, H = 0, Hpad = 0;
  double cell = 0, origin[3] = {0, 0, 0};
  bool sorted_positions = true;
  bool built = false;
  SceneDev host{};
  SceneDev* dev = nullptr;
  uint32_t* radix = nullptr;     // radix-sort block histograms and range bounds
  float4* positions = nullptr;   // [n] input positions (original order)
  uint32_t* perm = nullptr;      // [n] sorted indices
  uint32_t* skey = nullptr;      // [n] sorted keys
  float4* spos = nullptr;        // [n] sorted positions
  bool disposed = false;
};

struct pf_fluid_s {
  pf_particles_s* ps = nullptr;
  uint32_t start = 0, count = 0;
  double rest_density = 1000, spacing = 0, h = 0, compliance = 1e-4, mass = 0;
  uint32_t present = 0;
  double viscosity = 0, vorticity = 0, surface_tension = 0, adhesion = 0;
  struct Boundary { uint32_t start, count; bool dynamic; };
  std::vector<Boundary> boundaries;
  pf_loop_s* loop = nullptr;
  int slot = -1;  // index in SceneDev::fluid
  uint32_t* nl = nullptr;
  uint32_t* nl16 = nullptr;
  uint32_t tu_limit = PF_TU_MAX;
  uint32_t* tu_size = nullptr;
  uint32_t* tu_slots = nullptr;
  uint32_t* nl_count = nullptr;
  float4* normal = nullptr;
  float4* omega = nullptr;
  float4* dv = nullptr;
  int* reaction = nullptr;
  struct pf_viscosity_s* isv = nullptr;  // attached ViscositySolver (set by the loop)
  bool disposed = false;
};

struct pf_viscosity_s {
  pf_fluid_s* fluid = nullptr;
  double viscosity = 0;
  uint32_t iterations = 12;
  pf_loop_s* loop = nullptr;
  float4* rhs = nullptr;
  float4* a = nullptr;
  float4* b = nullptr;
  bool disposed = false;
};

struct pf_primitives_s {
  pf_particles_s* ps = nullptr;
  uint32_t reserved = 0;
  struct Prim {
    uint32_t kind, flags;
    double d0[4], d1[4], vel[4], spin[3], rot[4];
    bool has_pose = false;
    double prev_pos[3], prev_rot[4], offset[4], direction[3];
  };
  std::vector<Prim> prims;
  uint32_t capacity = 0;  // fixed once a loop is built
  pf_loop_s* loop = nullptr;
  uint32_t* packed = nullptr;
  float4 *d0 = nullptr, *d1 = nullptr, *vel = nullptr, *rot = nullptr, *spin = nullptr;
  bool dirty = true;
  bool disposed = false;
  void upload();
};

struct pf_sdf_collider_s {
  pf_particles_s* ps = nullptr;
  pf_loop_s* loop = nullptr;
  int slot = -1;                 // index in the scene's SdfDev table
  uint32_t res[3] = {0, 0, 0};
  double origin[3] = {0, 0, 0}, voxel[3] = {1, 1, 1};
  double pos[3] = {0, 0, 0}, quat[4] = {0, 0, 0, 1}, scale = 1, thickness = 0;
  double mus = 0.5, muk = 0.4;
  uint16_t* grid = nullptr;      // device half floats
  bool disposed = false;
  void write(SdfDev& D) const;   // the f32 uniforms
};

struct pf_softbody_s {
  pf_particles_s* ps = nullptr;
  struct Body { uint32_t start, count, surface_count; double compliance; };
  std::vector<Body> bodies;
  std::vector<float> rest;     // 4n (id order), w = compliance
  std::vector<float> weights;  // n
  uint32_t* d_start = nullptr;
  uint32_t* d_count = nullptr;
  float* d_compliance = nullptr;
  float4* d_rest = nullptr;
  float* d_weights = nullptr;
  float4* d_lambda = nullptr;
  float4* d_centers = nullptr;
  float4* d_rotations = nullptr;
  // local shape matching
  uint32_t mode = 0;             // PF_SHAPE_MATCHING_*
  uint32_t span_lo = 0, span_hi = 0;
  uint32_t* d_csr_off = nullptr;
  uint32_t* d_csr_idx = nullptr;
  float4* d_rest_center = nullptr;
  float4* d_group_center = nullptr;
  float4* d_group_rot = nullptr;
  float4* d_pair_lambda = nullptr;
  int* d_acc = nullptr;
  float4* d_q = nullptr;
  float4* d_qp = nullptr;
  float4* d_omega = nullptr;
  bool self_collision = false;   // selfCollision (SoftbodySystem.ts:71-72)
  pf_loop_s* loop = nullptr;
  bool disposed = false;
};

// A launch plan: a uniform substep schedule over one or more scenes.
struct PlanShape {
  uint32_t substeps = 4, iterations = 2;
  uint32_t nfluids = 0;
  uint32_t fx[PF_MAX_FLUIDS] = {0, 0, 0, 0};  // union of effects per fluid slot
  bool any_boundary = false;
  uint32_t nb[PF_MAX_FLUIDS] = {0, 0, 0, 0};    // boundary ranges per fluid slot (max over scenes)
  uint32_t bdyn[PF_MAX_FLUIDS] = {0, 0, 0, 0};  // union of dynamic-boundary bits per fluid slot
  uint32_t tu_cap[PF_MAX_FLUIDS] = {PF_TU_MAX, PF_TU_MAX, PF_TU_MAX, PF_TU_MAX};  // LDS union capacity (max over scenes)
  uint32_t isv[PF_MAX_FLUIDS] = {0, 0, 0, 0};  // viscosity-solver sweeps per fluid slot (max over scenes)
  bool local_soft = false;  // any scene's soft bodies use local shape matching
  bool walls = false;       // any scene has primitives
  bool soft = false;        // any scene has soft bodies
  bool grid = false;        // the scenes sort by hash bucket every substep (fluids or contacts)
  bool contact = false;     // any scene has particle contacts
  uint32_t grid_cap = 0;    // tile kernels' launch shape: 0 one workgroup per tile, else the persistent cap
};

struct pf_loop_s {
  pf_particles_s* ps = nullptr;
  pf::Context* ctx = nullptr;
  uint32_t substeps = 4, iterations = 2;
  double gravity[3] = {0, -9.81, 0};
  std::vector<pf_fluid_s*> fluids;
  std::vector<pf_viscosity_s*> viscosity;  // solvers, each directly after its fluid
  pf_softbody_s* soft = nullptr;
  pf_primitives_s* prims = nullptr;
  std::vector<pf_sdf_collider_s*> sdfs;
  pf_grid_s grid;  // the loop's grid view (owned_by_loop)
  Tile* tiles = nullptr;
  uint32_t ntiles = 0;
  uint2* stiles = nullptr;
  uint32_t nstiles = 0;
  uint2* btiles = nullptr;
  uint32_t nbtiles = 0;
  uint2* ltiles = nullptr;   // (scene, first id) tiles over the local soft-body span
  uint32_t nltiles = 0;
  Tile* ctiles = nullptr;    // (scene, first contact) tiles of the per-contact passes
  uint32_t nctiles = 0;
  bool contact = false;      // SimLoop contact option (loop.ts:23-49)
  uint32_t grid_cap = 0;     // launch shape of the tile kernels (pf_loop_set_launch_shape)
  bool static_boundaries = false;  // a fluid has a static boundary (volumes computed on the first step)
  double last_dt = -1;
  uint32_t last_flag_half = 0;
  uint64_t steps = 0;
  bool use_graph = false;
  hipGraphExec_t graph[4] = {nullptr, nullptr, nullptr, nullptr};  // by descriptor variant
  bool graph_valid[4] = {false, false, false, false};
  double graph_dt = -1;
  uint32_t kernels_per_step = 0;
  bool disposed = false;
};

// Scenes of a batch that share one substep schedule (their PlanShapes equal
// but for the union capacity and launch shape), enqueued over their own tile
// lists: a plain fluid scene keeps the fused schedule beside scenes whose soft
// bodies or contacts unfuse theirs.
struct BatchGroup {
  std::vector<uint32_t> scenes;  // indices into pf_batch_s::loops
  PlanShape shape;
  Tile* tiles = nullptr;
  uint32_t ntiles = 0;
  uint2* stiles = nullptr;
  uint32_t nstiles = 0;
  uint2* btiles = nullptr;
  uint32_t nbtiles = 0;
  uint2* ltiles = nullptr;
  uint32_t nltiles = 0;
  Tile* ctiles = nullptr;
  uint32_t nctiles = 0;
};

struct pf_batch_s {
  pf::Context* ctx = nullptr;
  std::vector<pf_loop_s*> loops;
  SceneDev* dev = nullptr;
  std::vector<BatchGroup> groups;
  void* scan_ctl = nullptr;
  PlanShape shape;  // every scene's merged; shape.grid_cap: launch shape of the tile kernels (pf_batch_set_launch_shape)
  bool use_graph = false;
  hipGraphExec_t graph = nullptr;
  bool graph_valid = false;
  bool desc_dirty = true;
};

namespace pf {
// The substep schedule, shared by loops and batches (sim.hip).
uint32_t enqueue_step(Context* ctx, const SceneDev* scenes, const PlanShape& shape,
                      const Tile* tiles, uint32_t ntiles, const uint2* stiles, uint32_t nstiles,
                      const uint2* btiles, uint32_t nbtiles, const uint2* ltiles, uint32_t nltiles,
                      const Tile* ctiles, uint32_t nctiles, void* scan_ctl);
int step_variant_flips(const PlanShape& sh);
// Multi-scene launches (pf_batch, pf_server; sim.hip).
PlanShape loop_plan_shape(const pf_loop_s* L);
void plan_merge(PlanShape& into, const PlanShape& scene);
void loop_append_tiles(const pf_loop_s* L, uint32_t scene, std::vector<Tile>& tiles, std::vector<uint2>& stiles,
                       std::vector<uint2>& btiles, std::vector<uint2>& ltiles, std::vector<Tile>& ctiles);
// The loop sorts its particles into a hash grid every substep.
bool loop_has_grid(const pf_loop_s* L);
bool loop_prepare_step(pf_loop_s* L, double dt);
void loop_finish_step(pf_loop_s* L, int flips);
}  // namespace pf
