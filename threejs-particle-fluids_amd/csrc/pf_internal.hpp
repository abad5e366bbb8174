// Internal device-side layout and math shared by every HIP kernel of the
// particle solver (gfx950 / CDNA4, wave64).
//
// Particle state lives in HBM as cell-sorted structure-of-arrays: after every
// substep's counting sort each particle sits at the slot of its hash bucket,
// and `ids[slot]` records which original particle it is. Kernels read scene
// parameters and buffer pointers from a `SceneDev` descriptor in device
// memory, selected per workgroup through a tile list, so one launch can
// advance one scene (plain form) or many independent scenes (batched form).
//
// Floating-point convention: every kernel is compiled with
// -ffp-contract=off and evaluates expressions in the operand order of the
// reference TSL passes (cited per function), with IEEE division and sqrt.
// The CPU oracle follows the same order, so single-substep results agree
// bit-for-bit when the neighbour order agrees (it does: stable sort by id).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define PF_MAX_FLUIDS 4
#define PF_MAX_NEIGHBORS 64          // neighborList.ts:14
#define PF_FLUID_MAX_PARTICLES (1u << 24)  // particles of a system with fluids (32-bit list offsets)
#define PF_MAX_BOUNDARIES 8
#define PF_MAX_SUBSTEPS 64
#define PF_MORTON_BIAS 512           // mortonHash.ts:25
#define PF_CELL_COORD_MAX (1 << 30)  // cellIndex.ts:35
#define PF_SCAN_TILE 4096            // buckets per scan workgroup (256 x 16)
#define PF_BLOCK 256
#define PF_STABLE_SEGMENT_LIMIT 2048 // buckets longer than this keep atomic order (flagged)
// Tile union: the particles of every bucket that the fluid particles of one
// 256-slot tile walk, staged in LDS for the neighbour passes.
#define PF_TU_MAX 2048               // particles a tile union may hold (16-bit list entries)
#define PF_TU_HASH 1024              // bucket hash-set slots per tile (power of two)
#define PF_TU_HASH_LIMIT 768         // distinct buckets beyond which the tile uses global gathers
#define PF_TU_GLOBAL 0xffffffffu     // tu_size value of a tile whose list holds global slots

// Overflow / status flag words, one u32 each in SceneDev::flags.
enum {
  PF_FLAG_GRID = 0,        // cell coordinate outside the Morton range (cellIndex.ts:150-169)
  PF_FLAG_NEIGHBORS = 1,   // a fluid neighbour list was truncated at 64 (neighborList.ts:63-66)
  PF_FLAG_POSITIONS = 2,   // position accumulator saturated (accumulator.ts:114-120)
  PF_FLAG_VELOCITIES = 3,  // velocity accumulator saturated
  PF_FLAG_IMPULSES = 4,    // surface-tension/adhesion accumulator saturated
  PF_FLAG_LONG_BUCKET = 5, // a bucket exceeded the stable-rank limit (order not canonical)
  PF_FLAG_REACTION = 6,    // boundary reaction accumulator saturated
  PF_NUM_FLAGS = 8
};

// Fluid effect bits (FluidSystem.ts:31-54: optional effects compiled in only when given).
enum {
  PF_FX_VISCOSITY = 1u << 0,
  PF_FX_VORTICITY = 1u << 1,
  PF_FX_SURFACE_TENSION = 1u << 2,
  PF_FX_ADHESION = 1u << 3,
};

// Primitive kinds / flags (primitives.ts:7-20).
enum { PF_KIND_PLANE = 0, PF_KIND_SPHERE = 1, PF_KIND_BOX = 2, PF_KIND_CAPSULE = 3 };
enum { PF_PRIM_INVERT = 1u << 0, PF_PRIM_SWEEP = 1u << 1 };

struct FluidDev {
  uint32_t start, count;   // particle id range of the fluid
  uint32_t effects;        // PF_FX_* bits
  uint32_t nboundary;      // boundary ranges (FluidSystem.addBoundary)
  uint32_t bstart[PF_MAX_BOUNDARIES], bcount[PF_MAX_BOUNDARIES];
  uint32_t bdynamic;       // bit b: boundary b recomputed every substep
  uint32_t rcover_lo, rcover_hi; // id span covering every boundary range (reaction apply)
  float rho0, pv, mass;    // rest density, spacing^3, rho0*spacing^3 (f32 uniforms)
  float inv_mass;          // 1 / mass, rounded once
  float h, h2, poly6, spiky; // kernels.ts:22-47 (coefficients computed in double)
  float compliance;
  float gamma, visc, vort, adhesion;
  float coh_coef, coh_h6_64;        // cohesion.ts:46-55 constants
  float adh_coef, adh_m4_over_h, adh_2h; // adhesion.ts:16-27 constants
  float imp_scale, imp_inv;         // Accumulator(50) for impulses
  float rea_scale, rea_inv;         // Accumulator(10) for boundary reaction
  uint32_t* nl;        // [64 * n] neighbour slots, column-major k*n + slot (global-gather tiles)
  uint32_t* nl16;      // [32 * n] pairs of 16-bit tile-union indices, column-major (k/2)*n + slot
  uint32_t tu_limit;   // largest union a tile stages (<= PF_TU_MAX)
  uint32_t* tu_size;   // [tiles] union size of each 256-slot tile, or PF_TU_GLOBAL
  uint32_t* tu_slots;  // [tiles * PF_TU_MAX] sorted slot of each union entry
  uint32_t* nl_count;  // [n]
  float4* normal;      // [2 n] pair records: (prediction, density), (colour-field normal, id bits)
  float4* omega;       // [n] vorticity (xyz) and |omega| (w)
  float4* dv;          // [n] XSPH velocity change
  int* reaction;       // [3n] fixed-point boundary reaction (only with boundaries)
  // ViscositySolver listed after this fluid (ViscositySolver.ts:38-123)
  uint32_t isv_iters;  // Jacobi sweeps per substep, 0: no solver
  float isv_nu;        // kinematic viscosity
  float isv_c10;       // 10 / h^2, computed in double
  float4* isv_rhs;     // [n] this substep's velocities, slot order
  float4* isv_a;       // [n] iterates (ping, pong); zero outside the fluid
  float4* isv_b;
};

// Global shape matching of one scene's soft bodies, in original-id order
// (SoftbodySystem.ts:105-199).
struct SoftDev {
  uint32_t nbodies;
  const uint32_t* body_start;
  const uint32_t* body_count;
  const float* body_compliance;
  const float4* rest;   // [n] rest offset from the body's rest centre, w = compliance
  const float* weights; // [n] mass weight relative to the body mean (massWeights)
  float4* lambda;       // [n] XPBD multiplier per particle
  float4* centers;      // [nbodies]
  float4* rotations;    // [3 * nbodies] rows of R
  // Local shape matching (shapeMatchImplicit.ts, rotation.ts), id order over
  // the span of every body.
  uint32_t local;       // 1: local shape matching
  uint32_t span_lo, span_hi;
  const uint32_t* csr_off;   // [n + 1] row offsets; row i = i, then its edge neighbours
  const uint32_t* csr_idx;   // [csr_off[n]]
  const float4* rest_center; // [n] rest centre of each particle's group
  float4* group_center;      // [n] current centre c_i
  float4* group_rot;         // [3n] rows of R_i
  float4* pair_lambda;       // [csr_off[n]] multiplier per (group, member)
  int* acc;                  // [3n] Accumulator(10) of the pulls
  float4* q;                 // [n] orientation (rotation)
  float4* qp;                // [n] predicted orientation
  float4* omega;             // [n] angular velocity
  float ai;                  // r^2 / 5 (orientation term weight)
  float maxc2;               // (5 r)^2 (pull clamp)
};

// A signed distance field collider (SDFCollider.ts:68-311): half-float
// voxels and the placement, all as the f32 uniforms the reference uploads.
struct SdfDev {
  const uint16_t* grid;           // [nx * ny * nz], x fastest
  int nx, ny, nz;
  float res[3], inv_res[3];       // resolution as float and 1 / resolution
  float origin[3];                // corner of the first voxel (local)
  float voxel[3], inv_voxel[3];
  float pos[3];                   // world position of the local origin
  float rot[9], inv_rot[9];       // row-major rotation and its transpose
  float scale, inv_scale;
  float thickness;
  float mus, muk;
};
#define PF_MAX_SDF 4

// Particle-particle contacts (src/core/contact, loop.ts:346-418). Records hold
// sorted slots of the substep that generated them; the multipliers are
// fixed point, 2^20 ticks per kg m (ContactBuffer.ts:16).
#define PF_MAX_QUIET 8
#define PF_CT_LAMBDA_SCALE 1048576.0f
struct ContactDev {
  uint32_t on;               // the scene's loop has contacts
  uint32_t max;              // maxContacts
  uint32_t blocks;           // workgroups of the per-contact passes (grid stride)
  uint32_t nquiet;           // extra noSelfContacts ranges (fluids are always quiet)
  uint32_t qstart[PF_MAX_QUIET], qcount[PF_MAX_QUIET];
  float dist, cr, cr2;       // 2r, 2r * 1.1 and (2r * 1.1)^2 (generate.ts:19)
  float mus, muk;
  uint32_t* count;           // [1] pairs emitted this substep, dropped ones included
  uint2* pair;               // [max] (emitter slot, other slot)
  int2* lam;                 // [max] (lambda_n, lambda_t) ticks
  float4* nrm;               // [max] last normal the solve used
  int4* acc;                 // [n] fixed-point sums (xyz) scattered by the per-contact passes and by
                             // surface tension's pairs (also without contacts); .w unused
};

struct SceneDev {
  uint32_t n;          // particle capacity
  uint32_t H;          // hash table size (power of two)
  uint32_t hash_mask;
  uint32_t nscan_tiles;
  float cell, ox, oy, oz;  // grid cell size and hash origin
  float radius;            // particle radius (collider contact distance)
  float dt;                // substep length (f32 uniform)
  float gx, gy, gz;        // gravity
  uint32_t nfluids;
  uint32_t has_boundary;   // any particle has boundaryVolume > 0
  // permuted particle state, double buffered by substep parity
  float4* x[2];
  float4* v[2];
  float* w[2];             // inverse mass
  uint32_t* id[2];         // original particle id at each slot
  float* rho[2];           // fluid density (persists: surface tension reads the previous one)
  float* bvol[2];          // boundary volume (Akinci), 0 for non-boundary particles
  uint32_t* grp[2];        // collision group
  // per-substep buffers
  float4* xs[2];           // predicted positions ping-pong; .w carries the fluid lambda
  float* mass;             // 1/w or 0, per slot
  uint32_t* slot_of;       // original id -> slot
  // counting sort
  uint32_t* key;           // [n] bucket per pre-sort slot
  uint32_t* rank;          // [n] atomic arrival rank inside the bucket
  uint32_t* counts;        // [H] histogram, left zeroed by the scan
  uint32_t* start;         // [H+1] exclusive prefix (cellStart); start[b+1] is cellEnd
  unsigned long long* scan_status; // [nscan_tiles] decoupled look-back words
  uint32_t* scan_ctl;      // [2] tiles finished this scan, scan epoch
  uint32_t* tmp_src;       // [n] pre-sort slot at each bucket position (arrival order)
  uint32_t* tmp_id;        // [n] original id at each bucket position (arrival order)
  uint32_t* tmp_key;       // [n] bucket at each bucket position
  uint32_t* flags;         // [PF_NUM_FLAGS]
  // colliders (one PrimitiveSet per scene)
  uint32_t nprim, prim_cap;
  const uint32_t* prim_packed; // (kind << 16) | flags
  const float4* prim_d0;
  const float4* prim_d1;       // .w = muS
  const float4* prim_vel;      // .w = muK
  const float4* prim_rot;      // quaternion xyzw
  const float4* prim_spin;     // angular velocity xyz
  float* prim_lambda;          // [2 * prim_cap * n] (lambda_n, lambda_t) per slot/primitive
  uint32_t* prim_touch;        // [n] bit p set when primitive p touched the slot this substep
  float pos_scale, pos_inv, vel_scale, vel_inv; // Accumulator(10), Accumulator(50)
  uint32_t nsdf;
  SdfDev sdf[PF_MAX_SDF];
  float* sdf_lambda;           // [2 * PF_MAX_SDF * n] (lambda_n, lambda_t) per slot/collider
  uint32_t* sdf_touch;         // [n] bit k set when collider k touched the slot this substep
  float clock[PF_MAX_SUBSTEPS];                 // sweep clock per substep (PrimitiveSet.ts:393-398)
  FluidDev fluid[PF_MAX_FLUIDS];
  SoftDev soft;
  ContactDev ct;
};


// One workgroup's share of a per-particle kernel: scene index and first slot.
struct Tile {
  uint32_t scene;
  uint32_t first;
};

// ---------------------------------------------------------------------------
// Integer helpers (bit-exact with the reference's WGSL semantics)
// ---------------------------------------------------------------------------

// Spread the low 10 bits of n to every third bit (mortonHash.ts:39-50).
__host__ __device__ __forceinline__ uint32_t pf_part1by2(uint32_t n) {
  n &= 0x000003ffu;
  n = (n | (n << 16)) & 0xff0000ffu;
  n = (n | (n << 8)) & 0x0300f00fu;
  n = (n | (n << 4)) & 0x030c30c3u;
  n = (n | (n << 2)) & 0x09249249u;
  return n;
}

// Unmasked Morton code of signed cell coordinates (mortonHash.ts:71-79).
// The bias add is done in unsigned arithmetic: two's-complement wrap, as WGSL.
__host__ __device__ __forceinline__ uint32_t pf_morton(int cx, int cy, int cz) {
  const uint32_t ux = (uint32_t)cx + (uint32_t)PF_MORTON_BIAS;
  const uint32_t uy = (uint32_t)cy + (uint32_t)PF_MORTON_BIAS;
  const uint32_t uz = (uint32_t)cz + (uint32_t)PF_MORTON_BIAS;
  return pf_part1by2(ux) | (pf_part1by2(uy) << 1) | (pf_part1by2(uz) << 2);
}

// Saturating f32 -> i32 (WGSL conversion semantics; NaN maps to 0).
__host__ __device__ __forceinline__ int pf_f2i_sat(float f) {
  if (!(f == f)) return 0;
  if (f >= 2147483648.0f) return 2147483647;
  if (f < -2147483648.0f) return (-2147483647 - 1);
  return (int)f;
}

__host__ __device__ __forceinline__ int pf_clampi(int v, int lo, int hi) {
  return v < lo ? lo : (v > hi ? hi : v);
}

// Cell coordinate for binning: FP32 subtract, then divide, floor, saturate,
// clamp to +-2^30 (cellIndex.ts:113-140). Returns 1 in *over when the clamp
// fired or the coordinate is outside the Morton range (cellIndex.ts:150-169).
__device__ __forceinline__ int pf_bin_coord(float p, float o, float cell, uint32_t* over) {
  const float rel = (p - o) / cell;
  const int raw = pf_f2i_sat(floorf(rel));
  const int c = pf_clampi(raw, -PF_CELL_COORD_MAX, PF_CELL_COORD_MAX - 1);
  if (raw != c || c < -PF_MORTON_BIAS || c >= PF_MORTON_BIAS) *over = 1u;
  return c;
}

// Query cell coordinate: floor + conversion without clamp (query.ts:30-33).
__device__ __forceinline__ int pf_query_coord(float p, float o, float cell) {
  return pf_f2i_sat(floorf((p - o) / cell));
}

// ---------------------------------------------------------------------------
// Correctly rounded reciprocal and square root in fewer instructions than the
// compiler's general IEEE sequences, over the ranges the passes use. Both
// start from the hardware approximation (v_rcp_f32, v_sqrt_f32: within 1 ulp)
// and round it exactly with fused multiply-adds, so the results equal IEEE
// 1.0f / x and sqrtf(q) bit for bit; pf_check_exact_math runs the exhaustive
// comparison on the device.
// ---------------------------------------------------------------------------

// 1/x for 2^-126 <= |x| < 2^126: one Newton-Raphson step from the hardware
// reciprocal y, y + y*(1 - x*y) with both operations fused (Markstein's
// iteration); 3 instructions against 10 for the general divide.
__device__ __forceinline__ float pf_rcp_rn(float x) {
  const float y = __builtin_amdgcn_rcpf(x);
  const float e = __builtin_fmaf(-x, y, 1.0f);
  return __builtin_fmaf(e, y, y);
}

// sqrt(q) for q = 0 and q >= 2^-96, and +inf: the hardware root s, moved to a
// neighbour when the exact residual q - s'*s says the neighbour rounds
// better (the rounding test of Markstein, IA-64 and Elementary Functions,
// 2000). 9 instructions.
__device__ __forceinline__ float pf_sqrt_rn_fast(float q) {
  const float s = __builtin_amdgcn_sqrtf(q);
  const float sm = __int_as_float(__float_as_int(s) - 1);
  const float sp = __int_as_float(__float_as_int(s) + 1);
  const float rm = __builtin_fmaf(-sm, s, q);
  const float rp = __builtin_fmaf(-sp, s, q);
  const float s1 = (rm <= 0.0f) ? sm : s;
  return (rp > 0.0f) ? sp : s1;
}

// sqrt(q) for every q >= 0 (and NaN): arguments below 2^-96 are scaled by
// 2^64 into the fast form's range and its root by 2^-32 back, both exact
// (powers of two, the scaled root normal), so the rounding is that of
// sqrt(q). Branch-free, 14 instructions against 16 for the general sequence.
__device__ __forceinline__ float pf_sqrt_rn(float q) {
  const bool tiny = q < 0x1p-96f;
  const float s = pf_sqrt_rn_fast(tiny ? q * 0x1p64f : q);
  return tiny ? s * 0x1p-32f : s;
}

// 0 < q < 2^-96, the one range of q >= 0 where pf_sqrt_rn_fast is not exact
// and pf_sqrt_rn scales. Passes test a group of square roots with it and take
// the direct form (DIRECT = true below) when no lane of the wave needs the
// scaled one: two instructions per root instead of five.
__device__ __forceinline__ bool pf_sqrt_scaled(float q) {
  return __float_as_uint(q) - 1u < 0x0F800000u - 1u;  // 0x0F800000 = 2^-96
}

// ---------------------------------------------------------------------------
// SPH kernels (kernels.ts:64-107), FP32, operand order as the reference
// ---------------------------------------------------------------------------

// W = max(h^2 - r^2, 0)^3 * poly6Coef, evaluated ((d*d)*d)*coef.
__device__ __forceinline__ float pf_poly6(float r2, float h2, float poly6) {
  const float d = fmaxf(h2 - r2, 0.0f);
  return ((d * d) * d) * poly6;
}

// grad W = r_vec * sp, sp = -(t*t*spikyCoef) / max(|r|, 1e-20), t = max(h - |r|, 0)
// (one rounded reciprocal: the reference's WGSL division is within 2.5 ulp).
// The passes fold their per-neighbour factor into sp before it meets r_vec,
// one product per component, as the oracle does. DIRECT: |r|^2 is known not
// to need the scaled square root (pf_sqrt_scaled false).
template <bool DIRECT = false>
__device__ __forceinline__ float pf_spiky_scale(float3 r, float h, float spiky) {
  const float rsq = (r.x * r.x + r.y * r.y) + r.z * r.z;
  const float rl = DIRECT ? pf_sqrt_rn_fast(rsq) : pf_sqrt_rn(rsq);
  const float t = fmaxf(h - rl, 0.0f);
  const float inv = pf_rcp_rn(fmaxf(rl, 1e-20f));
  const float mag = -((t * t) * spiky);
  return inv * mag;
}
__device__ __forceinline__ float3 pf_spiky_grad(float3 r, float h, float spiky) {
  const float sp = pf_spiky_scale(r, h, spiky);
  return make_float3(r.x * sp, r.y * sp, r.z * sp);
}

__device__ __forceinline__ float pf_dot3(float3 a, float3 b) {
  return (a.x * b.x + a.y * b.y) + a.z * b.z;
}
__device__ __forceinline__ float3 pf_sub3(float3 a, float3 b) {
  return make_float3(a.x - b.x, a.y - b.y, a.z - b.z);
}
__device__ __forceinline__ float3 pf_add3(float3 a, float3 b) {
  return make_float3(a.x + b.x, a.y + b.y, a.z + b.z);
}
__device__ __forceinline__ float3 pf_mul3(float3 a, float s) {
  return make_float3(a.x * s, a.y * s, a.z * s);
}
__device__ __forceinline__ float3 pf_cross3(float3 a, float3 b) {
  return make_float3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ float pf_len3(float3 a) { return pf_sqrt_rn(pf_dot3(a, a)); }
__device__ __forceinline__ float3 pf_xyz(float4 a) { return make_float3(a.x, a.y, a.z); }
// Component-wise copy. A whole-struct float4 assignment out of an unrolled
// register array lowers to a memcpy that leaves the array in scratch memory.
__device__ __forceinline__ float4 pf_f4(float4 a) { return make_float4(a.x, a.y, a.z, a.w); }

// Global-address-space view of a device pointer. Loads through it compile to
// global_load, counted by vmcnt alone. Loads through a generic pointer (any
// pointer read out of a descriptor) compile to flat_load, which counts in
// lgkmcnt as well, so every wait for an LDS read also waits for the global
// loads still in flight, such as a neighbour-list prefetch.
// (The host pass of a HIP compile parses device code too; it gets plain pointers.)
#if defined(__HIP_DEVICE_COMPILE__)
#define PF_G __attribute__((address_space(1)))
#else
#define PF_G
#endif
template <class T>
__device__ __forceinline__ PF_G T* pf_g(T* p) {
  return (PF_G T*)p;
}
// atomicAdd (relaxed, device scope) through a global-address-space pointer.
__device__ __forceinline__ void pf_atomic_add_g(PF_G int* a, int v) {
  __hip_atomic_fetch_add(a, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Fixed-point accumulator add (accumulator.ts:62-68): i32(round(v * scale)),
// round-half-to-even, saturating conversion. Sums wrap like WGSL atomicAdd.
// WGSL's saturating f32 -> i32 conversion (NaN to 0) in one instruction:
// v_cvt_i32_f32 clamps out-of-range inputs and maps NaN to 0, the same
// function as pf_f2i_sat, which compiles to three branches;
// pf_check_exact_math compares the two over every float.
__device__ __forceinline__ int pf_f2i_sat_hw(float f) {
  int r;
  asm("v_cvt_i32_f32 %0, %1" : "=v"(r) : "v"(f));
  return r;
}

__device__ __forceinline__ int pf_ticks(float v, float scale) {
  return pf_f2i_sat_hw(rintf(v * scale));
}
__device__ __forceinline__ int pf_wrap_add(int a, int b) {
  return (int)((uint32_t)a + (uint32_t)b);
}

// Raise a status flag word shared by every particle of a scene: one atomic
// per wave (from its lowest active lane), and none once the word reads set,
// so a flag raised by most particles does not serialise on one address.
__device__ __forceinline__ void pf_flag_raise(uint32_t* word) {
  const unsigned long long active = __ballot(1);
  if ((threadIdx.x & 63u) == (uint32_t)(__ffsll((unsigned long long)active) - 1)) {
    if (__atomic_load_n(word, __ATOMIC_RELAXED) == 0u) atomicOr(word, 1u);
  }
}

// XCD-aware mapping: gfx950 places consecutive workgroups on different XCDs
// round-robin. Give each XCD a contiguous run of tiles so neighbour reads of
// spatially adjacent (cell-sorted) particles hit the same XCD's L2.
__device__ __forceinline__ uint32_t pf_xcd_remap(uint32_t bid, uint32_t nblocks) {
  const uint32_t nxcd = 8;
  if (nblocks < 2 * nxcd) return bid;
  const uint32_t per = nblocks / nxcd;      // full runs
  const uint32_t full = per * nxcd;
  if (bid >= full) return bid;              // remainder keeps identity
  return (bid % nxcd) * per + bid / nxcd;
}

// The tiles of this workgroup, each passed to body as its index in the tile
// list (XCD-remapped). Every tile kernel is instantiated twice. PERSIST false
// is the launch of one workgroup per tile: body runs once. PERSIST true is the
// persistent form (fewer workgroups than tiles, a multiple of the 8 XCDs): the
// workgroup walks the list grid-stride, b, b + G, b + 2G, ..., tiles that keep
// b's XCD under the remap. Separate instantiations keep the loop's live
// ranges out of the one-tile form's register allocation. body returns to end
// its tile; a barrier between tiles keeps one tile's LDS staging from
// overwriting what a slower wave still reads, so no body may place a barrier
// after a return that a whole wave can take (every body here returns per
// lane, or block-uniformly).
template <bool PERSIST, class Body>
__device__ __forceinline__ void pf_for_tiles(uint32_t ntiles, Body&& body) {
  if constexpr (!PERSIST) {
    body(pf_xcd_remap(blockIdx.x, ntiles));
  } else {
    for (uint32_t vb = blockIdx.x; vb < ntiles; vb += gridDim.x) {
      body(pf_xcd_remap(vb, ntiles));
      if (vb + gridDim.x < ntiles) __syncthreads();
    }
  }
}
