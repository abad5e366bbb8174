/*
 * pf.h — C ABI of the MI355X-native particle solver (libpf.so).
 *
 * The ABI is the drop-in boundary for threejs-particle-fluids' per-substep
 * solver: each entry point replaces one reference interface, cited as
 * path:line in the reference repository. Plain pointers and sizes only: no
 * streams or torch types cross it, and no device pointers except the
 * buffers of pf_grid_sort_noalloc, whose callers hold their data on the GPU. Every call returns 0 on
 * success or a PF_E* code; pf_last_error() (thread-local) then holds the
 * message, worded as the reference's "<Owner>: <reason>" errors so a host
 * facade can rethrow it unchanged.
 *
 * Particle data crosses the ABI in original particle order and in the
 * reference's layouts: vec4 xyzw float32 per particle for snapshots
 * (particles.ts:27-38), xyz triples for uploads (ParticleInit, particles.ts:11-24).
 * Inside the library particles are kept cell-sorted; readbacks undo that.
 */
#ifndef PF_H
#define PF_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PF_OK 0
#define PF_EINVAL 1       /* invalid argument (reference throws synchronously) */
#define PF_EDEVICE 2      /* HIP runtime failure */
#define PF_ESTATE 3       /* object disposed / wrong order of calls */
#define PF_ENOMEM 4       /* device allocation failed */

typedef struct pf_context_s pf_context;
typedef struct pf_particles_s pf_particles;
typedef struct pf_grid_s pf_grid;
typedef struct pf_fluid_s pf_fluid;
typedef struct pf_primitives_s pf_primitives;
typedef struct pf_softbody_s pf_softbody;
typedef struct pf_sdf_collider_s pf_sdf_collider;
typedef struct pf_loop_s pf_loop;
typedef struct pf_batch_s pf_batch;
typedef struct pf_viscosity_s pf_viscosity;
typedef struct pf_fence_s pf_fence;
typedef struct pf_server_s pf_server;

/* ---- errors / version ---------------------------------------------------- */
const char* pf_last_error(void);
uint32_t pf_abi_version(void);

/* ---- context: one GPU, one HIP stream ------------------------------------
 * Replaces the WebGPU renderer the reference threads through every object
 * (createParticleRenderer, src/core/renderer.ts:16-33). */
int pf_context_create(int device, pf_context** out);
int pf_context_destroy(pf_context* ctx);
int pf_context_synchronize(pf_context* ctx);
/* Wall time of the stream between begin and end, from HIP events. */
int pf_context_timer_begin(pf_context* ctx);
int pf_context_timer_end(pf_context* ctx, float* ms);
/* Per-kernel timing: while on, every launch is bracketed by HIP events on
 * the library's stream; entries aggregate launches by kernel name. */
int pf_context_set_profiling(pf_context* ctx, int on);
int pf_context_profile_count(pf_context* ctx, uint32_t* n);
int pf_context_profile_entry(pf_context* ctx, uint32_t index, char* name, uint32_t name_cap,
                             uint32_t* launches, double* total_ms);
int pf_context_profile_reset(pf_context* ctx);

/* ---- ParticleSystem (src/core/particles.ts:52-258) ----------------------- */
typedef struct pf_snapshot {
  float* positions;           /* 4*capacity, nullable */
  float* predicted_positions; /* 4*capacity, nullable */
  float* velocities;          /* 4*capacity, nullable */
  float* inv_mass;            /* capacity, nullable */
  float* rotation;            /* 4*capacity, nullable */
  float* predicted_rotation;  /* 4*capacity, nullable */
  float* angular_velocity;    /* 4*capacity, nullable */
} pf_snapshot;

/* new ParticleSystem(renderer, capacity, particleRadius) (particles.ts:81-113) */
int pf_particles_create(pf_context* ctx, uint32_t capacity, double particle_radius,
                        pf_particles** out);
/* Like pf_particles_create, but every device buffer of the scene built on
 * these particles (grid, materials, colliders, loop) lives in one arena of
 * reserved device addresses, so a pf_server can move the scene to host memory
 * and back without changing any device address. No counterpart in the
 * reference (its buffers are WebGPU storage buffers). */
int pf_particles_create_evictable(pf_context* ctx, uint32_t capacity, double particle_radius,
                                  pf_particles** out);
/* Device bytes the scene's arena holds while resident (0 for particles not
 * created evictable). */
int pf_particles_device_bytes(pf_particles* ps, uint64_t* bytes);
/* dispose() (particles.ts:234-257); idempotent. */
int pf_particles_destroy(pf_particles* ps);
/* uploadParticles(data, start) (particles.ts:121-160): xyz triples; velocity
 * nullable (0), inv_mass nullable (1), collision_group nullable (0). */
int pf_particles_upload(pf_particles* ps, uint32_t start, uint32_t count, const float* positions,
                        const float* velocities, const float* inv_mass,
                        const uint32_t* collision_group);
/* setInvMass(range, invMass) (particles.ts:163-173) */
int pf_particles_set_inv_mass(pf_particles* ps, uint32_t start, uint32_t count, double inv_mass);
/* setCollisionGroup(range, group) (particles.ts:176-186) */
int pf_particles_set_collision_group(pf_particles* ps, uint32_t start, uint32_t count,
                                     uint32_t group);
/* readback() -> ParticleSnapshot (particles.ts:192-227) */
int pf_particles_readback(pf_particles* ps, const pf_snapshot* out);
/* the boundaryVolume buffer (particles.ts:101; written by
 * buildBoundaryVolumeKernel, boundaryVolume.ts:26-52), in particle order */
int pf_particles_readback_boundary_volume(pf_particles* ps, float* volume);
int pf_particles_info(pf_particles* ps, uint32_t* capacity, double* particle_radius);

/* ---- HashGrid (src/core/hashGrid/HashGrid.ts:80-286) ---------------------
 * hash_table_size 0 = min(nextPow2(2*capacity), 2^20) (HashGrid.ts:121-124).
 * Any power of two up to 2^30 is accepted (the reference caps at 2^20,
 * HashGrid.ts:136-142, because of its two-level scan). origin nullable. */
int pf_grid_create(pf_particles* ps, double cell_size, uint32_t hash_table_size,
                   const double origin[3], pf_grid** out);
/* A grid over its own position buffer (for sorting points that are not a
 * ParticleSystem); positions uploaded with pf_grid_upload_positions. */
int pf_grid_create_standalone(pf_context* ctx, uint32_t capacity, double cell_size,
                              uint32_t hash_table_size, const double origin[3], pf_grid** out);
int pf_grid_upload_positions(pf_grid* grid, const float* positions4, uint32_t count);
/* rebuild() (HashGrid.ts:219-222); sort_positions=0 skips the sorted copy of
 * the positions (sortedPositions.ts:31-39) when only the sort is wanted. */
int pf_grid_rebuild(pf_grid* grid);
int pf_grid_set_sorted_positions(pf_grid* grid, int enabled);
/* readback() -> HashGridSnapshot (HashGrid.ts:230-249). counts/cell_start/
 * cell_end hold hashTableSizePadded entries; each pointer nullable. */
int pf_grid_readback(pf_grid* grid, uint32_t* cell_index, uint32_t* counts, uint32_t* cell_start,
                     uint32_t* cell_end, uint32_t* sorted_indices);
int pf_grid_readback_sorted_keys(pf_grid* grid, uint32_t* keys);
int pf_grid_readback_sorted_positions(pf_grid* grid, float* positions4);
/* readbackOverflow() (HashGrid.ts:256-260) */
int pf_grid_readback_overflow(pf_grid* grid, int* overflow);
int pf_grid_info(pf_grid* grid, uint32_t* capacity, uint32_t* hash_table_size,
                 uint32_t* hash_table_size_padded);
/* hashOrigin setter (HashGrid.ts:205-215) */
int pf_grid_set_origin(pf_grid* grid, const double origin[3]);
/* emitForEachNeighbor (query.ts:24-70) over a built standalone grid, as the
 * reference's hash-grid tests use it: per query point, the number of
 * candidates visited (own index excluded) and how many lie within radius. */
int pf_grid_query_counts(pf_grid* grid, const float* query4, uint32_t nquery, double radius,
                         uint32_t* candidates, uint32_t* within);
int pf_grid_destroy(pf_grid* grid);

/* No counterpart in the reference. For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops): the standalone
 * HashGrid rebuild + readback (pf_grid_create_standalone, _upload_positions,
 * _rebuild, _readback, _readback_sorted_keys, _readback_overflow) as one call
 * that allocates nothing. Every pointer but origin is DEVICE memory on the
 * context's GPU: positions4 [n] vec4 (16-byte aligned); outputs cell_index,
 * sorted_indices, sorted_keys [n] and counts, cell_start, cell_end [Hpad]
 * (hashTableSizePadded); overflow [1] nullable. The work is enqueued on the
 * context's stream and the call returns without waiting
 * (pf_context_synchronize waits). Its scratch is one fixed device array of
 * PF_NOALLOC_SCRATCH_BYTES per GPU, held until the process ends and shared by
 * every call on that GPU: calls that overlap on one GPU race. Sizes above
 * PF_NOALLOC_MAX_PARTICLES positions or PF_NOALLOC_MAX_BUCKETS buckets fail
 * with PF_EINVAL. Results equal the standalone grid's bit for bit. */
#define PF_NOALLOC_MAX_PARTICLES (1u << 22)
#define PF_NOALLOC_MAX_BUCKETS (1u << 22)
#define PF_NOALLOC_SCRATCH_BYTES 51389184u
int pf_grid_sort_noalloc(pf_context* ctx, const float* positions4, uint32_t n, double cell_size,
                         uint32_t hash_table_size, const double origin[3], uint32_t* cell_index,
                         uint32_t* counts, uint32_t* cell_start, uint32_t* cell_end,
                         uint32_t* sorted_indices, uint32_t* sorted_keys, uint32_t* overflow);

/* ---- Device self-check (no counterpart in the reference) -------------------
 * The passes take reciprocals and square roots through short sequences that
 * start from the hardware approximations and round them exactly, and convert
 * fixed-point ticks with the hardware's saturating conversion. This call
 * compares one of them bit for bit with its reference form (1.0f / x, sqrtf,
 * the branchy saturating conversion) for every float whose bit pattern lies
 * in [first_bits, end_bits), on the
 * context's GPU, and returns the number of mismatches and the first one
 * (0xffffffff if none). It is for tests and for users who move the library
 * to another GPU. Blocks until done. */
#define PF_EXACT_RCP 0        /* 1/x, exact for 2^-126 <= |x| < 2^126 */
#define PF_EXACT_SQRT 1       /* sqrt(q), exact for every q >= 0 and -0 */
#define PF_EXACT_SQRT_FAST 2  /* its branch-free part, exact for q = 0 and q >= 2^-96 */
#define PF_EXACT_F2I 3        /* saturating f32 -> i32 (NaN to 0), one instruction, for every x */
int pf_check_exact_math(pf_context* ctx, int function, uint32_t first_bits, uint32_t end_bits,
                        uint64_t* mismatches, uint32_t* first_mismatch);

/* ---- FluidSystem (src/fluids/FluidSystem.ts:71-322) ---------------------- */
#define PF_FLUID_VISCOSITY 1u
#define PF_FLUID_VORTICITY 2u
#define PF_FLUID_SURFACE_TENSION 4u
#define PF_FLUID_ADHESION 8u
typedef struct pf_fluid_options {
  uint32_t range_start, range_count; /* range_count 0 = every particle */
  double rest_density;               /* <= 0: default 1000 */
  double particle_spacing;           /* <= 0: default 2 * particle radius */
  double smoothing_radius;           /* <= 0: default 2 * spacing */
  double compliance;                 /* < 0: default 1e-4 */
  uint32_t present;                  /* PF_FLUID_* bits: effects given in the options */
  double viscosity, vorticity, surface_tension, adhesion;
} pf_fluid_options;

int pf_fluid_create(pf_particles* ps, const pf_fluid_options* options, pf_fluid** out);
/* viscosity / vorticity / surfaceTension / adhesion setters (FluidSystem.ts:129-159) */
int pf_fluid_set(pf_fluid* fluid, uint32_t effect, double value);
int pf_fluid_get(pf_fluid* fluid, uint32_t effect, double* value);
/* addBoundary(range, {dynamic}) (FluidSystem.ts:171-183) */
int pf_fluid_add_boundary(pf_fluid* fluid, uint32_t start, uint32_t count, int dynamic);
/* density buffer in original particle order (FluidSystem.ts:81) */
int pf_fluid_readback_density(pf_fluid* fluid, float* density);
/* readbackOverflow() (FluidSystem.ts:190-192) */
int pf_fluid_readback_overflow(pf_fluid* fluid, int* overflow);
/* mass, smoothing radius, particle spacing, rest density (FluidSystem.ts:75-84) */
int pf_fluid_info(pf_fluid* fluid, double* mass, double* smoothing_radius,
                  double* particle_spacing, double* rest_density);
/* No counterpart in the reference (a tuning knob of this implementation):
 * the most particles a 256-slot tile may stage in LDS for the neighbour
 * passes (at most 2048, the default). Tiles whose neighbour buckets hold
 * more gather from global memory; 0 makes every tile do so. Results do not
 * depend on it. */
int pf_fluid_set_tile_union_limit(pf_fluid* fluid, uint32_t max_particles);
/* No counterpart in the reference (a diagnostic of the same knob): after the
 * last neighbour build, the number of 256-slot tiles, how many gathered from
 * global memory, the largest staged union and the mean staged union. Each
 * output nullable. Waits for the loop's work. */
int pf_fluid_tile_union_stats(pf_fluid* fluid, uint32_t* tiles, uint32_t* global_tiles,
                              uint32_t* largest_union, double* mean_union);
int pf_fluid_destroy(pf_fluid* fluid);

/* ---- ViscositySolver (src/fluids/ViscositySolver.ts:38-131) -----------
 * Implicit viscosity for thick liquids: each substep solves
 * (I - dt nu L) v = v0 for the fluid's velocities with `iterations` Jacobi
 * sweeps over the fluid's neighbour lists. List it directly after its fluid
 * in pf_loop_create's materials (kind PF_MATERIAL_VISCOSITY). viscosity >= 0
 * (m^2/s, about 20 for honey); iterations 1..64 (default 12). */
int pf_viscosity_create(pf_fluid* fluid, double viscosity, uint32_t iterations, pf_viscosity** out);
/* the `viscosity` setter (ViscositySolver.ts:65-70) */
int pf_viscosity_set(pf_viscosity* solver, double viscosity);
int pf_viscosity_destroy(pf_viscosity* solver);

/* ---- Surface field (src/fluids/render/SurfaceField.ts:101-369) -----------
 * The liquid's surface as a voxel grid, the field FluidSurfaceRenderer
 * sphere-traces (FluidSurfaceRenderer.ts:181-230, update at 632-658),
 * computed headless from the committed particles: a cloud-in-cell splat,
 * three Gaussian passes and phi = |p - mean| - radius. Each voxel holds four
 * floats (phi in voxels, volume fraction, wall proximity, 1), x fastest.
 * Options < 0 (or 0 for voxel_budget) take the reference's defaults:
 * defaultVoxelBudget (FluidSurfaceRenderer.ts:590-592), radius 1.25 r,
 * meniscus height 1.6 r and width 2.5 r for particle radius r.
 * wet = 1 makes every primitive and SDF collider of the fluid's loop a
 * wetted wall (the `colliders` option): the liquid draws a meniscus there
 * and is carved out of them. The `solids` range's particles wet the liquid
 * too. Rendering options (appearance, carve, cavities) have no counterpart. */
typedef struct pf_surface_field_options {
  double min[3], max[3];  /* bounds (a Box3) */
  double voxel_budget;
  double motion_stretch;  /* seconds, >= 0 */
  uint32_t solids_start, solids_count;
  int wet;
  double radius, meniscus_height, meniscus_width;
} pf_surface_field_options;
typedef struct pf_surface_field_s pf_surface_field;
int pf_surface_field_create(pf_fluid* fluid, const pf_surface_field_options* options, pf_surface_field** out);
/* grid resolution, origin (bounds.min) and voxel edge in metres */
int pf_surface_field_info(pf_surface_field* field, uint32_t resolution[3], double origin[3], double* voxel);
/* update(): enqueues the field's passes on the fluid's stream */
int pf_surface_field_update(pf_surface_field* field);
/* copies resolution[0]*resolution[1]*resolution[2] float4 voxels to rgba */
int pf_surface_field_readback(pf_surface_field* field, float* rgba);
int pf_surface_field_destroy(pf_surface_field* field);

/* ---- PrimitiveSet (src/core/collision/PrimitiveSet.ts:121-518) -----------
 * mu_s / mu_k < 0 take the defaults 0.5 / 0.4 (PrimitiveSet.ts:529-544). */
int pf_primitives_create(pf_particles* ps, uint32_t capacity, pf_primitives** out);
int pf_primitives_add_plane(pf_primitives* set, const double normal[3], const double point[3],
                            double mu_s, double mu_k, uint32_t* slot);
int pf_primitives_add_sphere(pf_primitives* set, const double center[3], double radius,
                             int invert, double mu_s, double mu_k, uint32_t* slot);
/* rotation: quaternion xyzw, nullable (identity) */
int pf_primitives_add_box(pf_primitives* set, const double center[3], const double half_extents[3],
                          const double rotation[4], int invert, double mu_s, double mu_k,
                          uint32_t* slot);
int pf_primitives_add_capsule(pf_primitives* set, const double a[3], const double b[3],
                              double radius, double mu_s, double mu_k, uint32_t* slot);
/* setSphere(slot, center, radius, velocity?) (PrimitiveSet.ts:241-256) */
int pf_primitives_set_sphere(pf_primitives* set, uint32_t slot, const double center[3],
                             double radius, const double velocity[3]);
/* attach()+update() for a host-driven pose: place the primitive's reference
 * point at `position` with orientation `rotation` (xyzw); with dt > 0 the
 * surface velocity and spin follow from the previous pose
 * (PrimitiveSet.ts:211-285). The first call only records the pose. */
int pf_primitives_set_pose(pf_primitives* set, uint32_t slot, const double position[3],
                           const double rotation[4], double dt);
int pf_primitives_count(pf_primitives* set, uint32_t* count);
int pf_primitives_destroy(pf_primitives* set);

/* ---- SoftbodySystem (src/softbody/SoftbodySystem.ts:105-359) -------------
 * SoftbodyDef (SoftbodySystem.ts:28-63) and the shapeMatching option:
 * global (shapeMatch.ts) or local (shapeMatchImplicit.ts, rotation.ts). */
typedef struct pf_softbody_def {
  uint32_t start, count;   /* particle range of the body */
  uint32_t surface_count;  /* 0xffffffff = all */
  double compliance;       /* s^2/kg, >= 0 */
  const float* rest_positions; /* xyz triples, nullable = uploaded positions */
  const uint32_t* edges;   /* pairs [i0, j0, i1, j1, ...] local to the body, nullable */
  uint32_t edge_count;     /* number of pairs */
  uint32_t rest_count;     /* floats in rest_positions; must be 3 * count when it is given
                              (SoftbodySystem.ts:171-173) */
} pf_softbody_def;
#define PF_SHAPE_MATCHING_GLOBAL 0u
#define PF_SHAPE_MATCHING_LOCAL 1u
int pf_softbody_create(pf_particles* ps, const pf_softbody_def* bodies, uint32_t nbodies,
                       uint32_t shape_matching, pf_softbody** out);
/* selfCollision option (SoftbodySystem.ts:71-72, 218-233); before the loop
 * is built. Without it the SimLoop gives each body its own collision group
 * (unless the whole body already shares a non-zero one). */
int pf_softbody_set_self_collision(pf_softbody* sb, int enabled);
/* setCompliance(index, compliance) (SoftbodySystem.ts:220-224) */
int pf_softbody_set_compliance(pf_softbody* sb, uint32_t index, double compliance);
/* bodyCenters (vec4 per body) and bodyRotations (3 vec4 rows per body) */
int pf_softbody_readback_frames(pf_softbody* sb, float* centers4, float* rotations12);
int pf_softbody_destroy(pf_softbody* sb);

/* ---- SDFCollider (src/core/collision/SDFCollider.ts:68-311) --------------
 * data: one distance per voxel, x fastest, then y, then z (SDFData, 32-39);
 * stored as half floats converted like three's DataUtils.toHalfFloat
 * (clamped to +-65504, mantissa truncated) and sampled trilinearly with
 * clamp-to-edge addressing (sdf.ts:107-146). mu_s / mu_k < 0 take the
 * defaults 0.5 / 0.4; rotation (xyzw) is normalised, nullable = identity. */
typedef struct pf_sdf_options {
  double mu_s, mu_k;
  double position[3];
  const double* rotation;
  double scale;      /* <= 0: 1 */
  double thickness;  /* extra contact distance beyond the particle radius */
} pf_sdf_options;
int pf_sdf_collider_create(pf_particles* ps, const float* data, const uint32_t resolution[3],
                           const double origin[3], const double voxel_size[3],
                           const pf_sdf_options* options, pf_sdf_collider** out);
/* setPosition / setRotation / setScale (SDFCollider.ts:186-212) */
int pf_sdf_collider_set_position(pf_sdf_collider* c, const double position[3]);
int pf_sdf_collider_set_rotation(pf_sdf_collider* c, const double rotation[4]);
int pf_sdf_collider_set_scale(pf_sdf_collider* c, double scale);
/* muS / muK / thickness setters (SDFCollider.ts:150-172) */
int pf_sdf_collider_set_friction(pf_sdf_collider* c, double mu_s, double mu_k);
int pf_sdf_collider_set_thickness(pf_sdf_collider* c, double thickness);
int pf_sdf_collider_destroy(pf_sdf_collider* c);

/* ---- SDF binary format (src/sdf/writeBinary.ts:21-106) ------------------
 * decodeSdfBinary: validates the 48-byte PSDF header as the reference does
 * (same messages) and returns resolution, origin and voxel size (f32 in the
 * file, widened); the voxels (x fastest) are copied only when `data` holds
 * all `voxel_count` of them, so a first call with data = NULL reads the
 * header. encodeSdfBinary: writes the format; out = NULL returns the size. */
int pf_sdf_decode(const void* bytes, size_t size, uint32_t resolution[3], double origin[3], double voxel_size[3],
                  float* data, size_t data_capacity, size_t* voxel_count);
int pf_sdf_encode(const float* data, size_t data_count, const uint32_t resolution[3], const double origin[3],
                  const double voxel_size[3], void* out, size_t capacity, size_t* size);

/* ---- SimLoop (src/core/loop.ts:92-324) ----------------------------------- */
#define PF_MATERIAL_FLUID 1u
#define PF_MATERIAL_SOFTBODY 2u
#define PF_MATERIAL_VISCOSITY 3u
typedef struct pf_material {
  uint32_t kind;   /* PF_MATERIAL_* */
  void* handle;    /* pf_fluid*, pf_softbody* or pf_viscosity* */
} pf_material;
#define PF_COLLIDER_PRIMITIVES 1u
#define PF_COLLIDER_SDF 2u
typedef struct pf_collider {
  uint32_t kind;   /* PF_COLLIDER_* */
  void* handle;    /* pf_primitives* or pf_sdf_collider* */
} pf_collider;
typedef struct pf_loop_options {
  uint32_t substeps;      /* 0 = default 4 */
  uint32_t iterations;    /* 0xffffffff = default 2 */
  double gravity[3];
  int has_hash_origin;
  double hash_origin[3];
  /* contact: true | ContactOptions (loop.ts:23-31, 43-49): particle-particle
   * contacts. mu < 0 takes the defaults 0.5 / 0.4; max_contacts 0 takes
   * 8 x capacity. Fluids never contact their own particles; quiet adds up to
   * 8 more [start, count] ranges that never contact each other (a material's
   * noSelfContacts, materials.ts:45). */
  int contact;
  double contact_mu_s, contact_mu_k;
  uint32_t max_contacts;
  uint32_t nquiet;
  const uint32_t* quiet;
} pf_loop_options;
typedef struct pf_overflow {
  int positions, velocities, grid, contacts, neighbors, impulses, long_bucket;
} pf_overflow;

/* new SimLoop(particles, {substeps, iterations, gravity, materials, colliders})
 * (loop.ts:117-241): at most one PrimitiveSet and four SDFColliders. */
int pf_loop_create(pf_particles* ps, const pf_loop_options* options, const pf_material* materials,
                   uint32_t nmaterials, const pf_collider* colliders, uint32_t ncolliders,
                   pf_loop** out);
/* step(dt) (loop.ts:267-279): enqueues every substep and returns. */
int pf_loop_step(pf_loop* loop, double dt);
int pf_loop_synchronize(pf_loop* loop);
/* The completion behind step()'s Promise (loop.ts:267-279, which resolves
 * when computeAsync's work is done): a fence records the point of the
 * loop's stream after the steps enqueued so far; pf_fence_wait blocks until
 * the GPU passes it, pf_fence_query polls. Waiting touches neither the loop
 * nor its stream, so another thread may wait while the owner keeps stepping.
 * Destroy each fence once. */
int pf_loop_fence(pf_loop* loop, pf_fence** out);
/* The same behind everything enqueued on the context's stream (the batched
 * form's step Promise). */
int pf_context_fence(pf_context* ctx, pf_fence** out);
int pf_fence_wait(pf_fence* fence);
int pf_fence_query(pf_fence* fence, int* done);
int pf_fence_destroy(pf_fence* fence);
/* readbackOverflow() (loop.ts:286-295) */
int pf_loop_readback_overflow(pf_loop* loop, pf_overflow* out);
int pf_loop_set_substeps(pf_loop* loop, uint32_t substeps);
int pf_loop_set_gravity(pf_loop* loop, const double gravity[3]);
int pf_loop_info(pf_loop* loop, uint32_t* substeps, uint32_t* iterations, double* cell_size,
                 uint32_t* hash_table_size, uint32_t* kernels_per_step);
/* SimLoop.contacts.readbackCount / readbackPairs (ContactBuffer.ts:74-97):
 * pairs emitted in the last substep (dropped ones included) and up to
 * `cap` stored (i, j) particle-index pairs, i the particle that found the
 * pair. Fails when the loop has no contacts. */
int pf_loop_readback_contacts(pf_loop* loop, uint32_t* count, uint32_t* pairs, uint32_t cap,
                              uint32_t* stored);
/* The loop's grid (SimLoop.hashGrid, loop.ts:104-106); owned by the loop. */
int pf_loop_grid(pf_loop* loop, pf_grid** out);
/* Capture each step's launches in a HIP graph and replay it (1) or launch
 * them one by one (0, default). */
int pf_loop_set_graph(pf_loop* loop, int enabled);
int pf_loop_destroy(pf_loop* loop);

/* ---- Batched form: many independent loops advanced by one launch sequence.
 * No counterpart in the reference. All loops must share substeps and
 * iterations; every launch covers the work tiles of every scene. */
int pf_batch_create(pf_context* ctx, pf_loop* const* loops, uint32_t nloops, pf_batch** out);
int pf_batch_step(pf_batch* batch, double dt);
int pf_batch_set_graph(pf_batch* batch, int enabled);
int pf_batch_destroy(pf_batch* batch);

/* ---- Autotuner (no counterpart in the reference): times `steps` steps of
 * the loop's or batch's real workload for each candidate tile-union LDS
 * capacity (256..2048 particles staged per 256-slot tile; larger unions walk
 * global memory) and sets the fastest on every fluid. The tuning steps
 * advance the scenes. ms_per_step (n, nullable) receives each candidate's
 * mean over a forward and a backward sweep. */
int pf_loop_autotune(pf_loop* loop, double dt, const uint32_t* candidates, uint32_t n, uint32_t steps,
                     uint32_t* best, float* ms_per_step);
int pf_batch_autotune(pf_batch* batch, double dt, const uint32_t* candidates, uint32_t n, uint32_t steps,
                      uint32_t* best, float* ms_per_step);

/* ---- Launch shape (no counterpart in the reference; the reference's passes
 * are one compute dispatch each, src/core/loop.ts:276-278). workgroups = 0
 * launches every tile kernel of a step with one 256-thread workgroup per
 * 256-slot tile; a positive value (a multiple of 8, at most 2^20) is the
 * persistent form: each launch has at most that many workgroups, and each
 * walks the tile list grid-stride. Results do not depend on the shape. A
 * loop's or batch's graph is re-captured; a server applies it to launches
 * formed after the call. */
int pf_loop_set_launch_shape(pf_loop* loop, uint32_t workgroups);
int pf_batch_set_launch_shape(pf_batch* batch, uint32_t workgroups);
int pf_server_set_launch_shape(pf_server* srv, uint32_t workgroups);
/* The autotuner over launch shapes: times each candidate (0 or multiples of
 * 8) like pf_loop_autotune / pf_batch_autotune and sets the fastest. */
int pf_loop_autotune_launch(pf_loop* loop, double dt, const uint32_t* candidates, uint32_t n, uint32_t steps,
                            uint32_t* best, float* ms_per_step);
int pf_batch_autotune_launch(pf_batch* batch, double dt, const uint32_t* candidates, uint32_t n, uint32_t steps,
                             uint32_t* best, float* ms_per_step);

/* ---- Served form: a request loop over live scenes. No counterpart in the
 * reference. Step requests for scenes added to the server are queued as they
 * arrive; a worker thread launches every queued request whose scene is not
 * already on the GPU as one multi-scene launch (continuous batching) on one of
 * `slots` streams, so launches overlap on the device. Scenes stay resident in
 * HBM; admission is capped by `hbm_budget_bytes` of scene arenas: before a
 * launch, idle scenes are evicted to host memory, least recently stepped
 * first, to make room for the evicted scenes it needs back. A scene's requests
 * run in submission order, one step each. While a scene is served, the
 * ordinary calls on it (step, readback, upload, new materials) fail with
 * PF_ESTATE; pf_server_remove_scene hands it back, resident. */
typedef struct pf_server_options {
  uint64_t hbm_budget_bytes;     /* resident scene bytes; 0 = no cap */
  uint32_t max_batch_scenes;     /* scenes per launch; 0 = 256 */
  uint64_t max_batch_particles;  /* particles per launch; 0 = no cap */
  uint32_t slots;                /* launches in flight at once; 0 = 2 */
  uint64_t host_budget_bytes;    /* host bytes evicted scenes may hold; 0 = no cap. An
                                    eviction beyond it is refused: the scene stays resident */
} pf_server_options;
typedef struct pf_server_stats {
  uint64_t submitted, completed, failed;
  uint64_t launches;             /* multi-scene launches */
  uint64_t scene_steps;          /* completed requests' steps */
  uint64_t particle_substeps;    /* sum over completed requests of particles x substeps */
  uint64_t evictions, restores;
  uint64_t resident_bytes, peak_resident_bytes, hbm_budget_bytes;
  uint32_t scenes, resident_scenes;
  uint32_t max_launch_scenes;    /* largest launch so far */
  uint64_t host_bytes;           /* host memory evicted scenes hold now */
  uint64_t evict_failures;       /* evictions refused (host budget, host memory) */
} pf_server_stats;
/* Times of one request in seconds on the server's clock: submitted,
 * launched, finished (0 while not reached). status: 0 pending, 1 launched,
 * 2 done, 3 failed. */
typedef struct pf_request_info {
  uint32_t scene, status, launch_scenes, pad;
  double t_submit, t_launch, t_done;
} pf_request_info;
int pf_server_create(pf_context* ctx, const pf_server_options* opts, pf_server** out);
/* Takes a loop built on evictable particles; every served loop shares the
 * first one's substeps and iterations. Admission may evict idle scenes. */
int pf_server_add_scene(pf_server* srv, pf_loop* loop, uint32_t* scene_id);
/* Waits for the scene's queued requests, restores it if evicted and hands
 * it back to its owner. */
int pf_server_remove_scene(pf_server* srv, uint32_t scene_id);
/* Queues one step(dt) of a scene and returns at once; thread-safe. */
int pf_server_submit(pf_server* srv, uint32_t scene_id, double dt, uint64_t* ticket);
/* Blocks until the request finished (or failed). */
int pf_server_wait(pf_server* srv, uint64_t ticket, pf_request_info* info);
/* Blocks until every submitted request finished. */
int pf_server_drain(pf_server* srv);
int pf_server_request_info(pf_server* srv, uint64_t first_ticket, uint32_t count, pf_request_info* out);
int pf_server_stats_get(pf_server* srv, pf_server_stats* out);
/* Current server clock, seconds (the clock of pf_request_info). */
double pf_server_now(pf_server* srv);
/* Moves an idle scene to host memory now (also done by admission). */
int pf_server_evict_scene(pf_server* srv, uint32_t scene_id);
/* Drains, stops the worker and hands every scene back, resident. */
int pf_server_destroy(pf_server* srv);

/* ---- Scene construction (src/simulation/layout.ts:7-42, Simulation.ts:957-1104) */
/* fillBox(box, spacing): lattice points, y -> z -> x order. out may be null to count. */
int pf_layout_fill_box(const double box_min[3], const double box_max[3], double spacing,
                       float* out_xyz, uint32_t capacity, uint32_t* count);
/* radiusForBudget(volume, area, clothSpacing, budget) (layout.ts:27-42) */
double pf_layout_radius_for_budget(double volume, double area, double cloth_spacing,
                                   uint32_t budget);
/* Simulation.fitBudget for box fluids in a container (Simulation.ts:957-1002 with
 * layout 1044-1104): radius and particle count of the best fitting layout. */
int pf_layout_fit_budget(uint32_t budget, const double container[6], const double* fluid_boxes,
                         uint32_t nboxes, double* radius, uint32_t* count);

#ifdef __cplusplus
}
#endif
#endif /* PF_H */
