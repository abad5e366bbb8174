// Surface field of a particle liquid (src/fluids/render/SurfaceField.ts:
// 101-369), the rendering step after the solver, computed headless:
//   1. cloud-in-cell splat of every fluid particle (and of the wetted solid
//      range) into a voxel grid, in 16-bit fixed point with integer atomics:
//      mass and the mass-weighted offset from each voxel centre;
//   2. a separable Gaussian (three 1D passes) that re-bases the offsets onto
//      each destination voxel, so the result is the kernel-weighted mean
//      particle position around every voxel;
//   3. phi = |p - mean| - R (Zhu & Bridson 2005), lowered near wetted walls
//      into a meniscus and carved out of the colliders.
// All passes are O(voxels) or O(particles) streams over HBM; voxels are
// x-fastest, so a wave's taps along y and z are coalesced rows. The output
// is (phi in voxels, volume fraction, wall proximity, 1) per voxel in FP32;
// the reference stores the same values in a half-float 3D texture.
#include <cmath>

#include "collide.hpp"
#include "pf_internal.hpp"
#include "runtime.hpp"

using namespace pf;

namespace {

constexpr float kFixed = 65536.0f;  // FIXED (SurfaceField.ts:32)
constexpr float kBand = 6.0f;       // FIELD_BAND (SurfaceField.ts:34)

// The field's constants, as the reference turns its JS doubles into the f32
// literals and uniforms its kernels read, and the grid's buffers.
struct FieldDev {
  uint32_t n[3], count, taps, samples;
  uint32_t fluid_start, fluid_count, solid_start, solid_count, wet;
  float origin[3], dims[3], voxel, inv_voxel, nxny, nx;
  float weights[13];
  float stretch, reach3, inv_samples, vol_scale, solid_scale;
  float radius, men_h, men_w;
  int* mass;
  int* off;     // [3 * count]
  int* solid;
  float4* a;
  float4* b;
  float* sa;
  float* sb;
  float2* walls;  // (wetting proximity, nearest collider distance)
  float4* out;
};

__device__ __forceinline__ bool fd_in(uint32_t id, uint32_t start, uint32_t count) {
  return id >= start && id - start < count;
}

__global__ __launch_bounds__(PF_BLOCK) void k_field_clear(FieldDev F) {
  const uint32_t i = blockIdx.x * PF_BLOCK + threadIdx.x;
  if (i >= F.count) return;
  F.mass[i] = 0;
  F.off[3 * (size_t)i + 0] = 0;
  F.off[3 * (size_t)i + 1] = 0;
  F.off[3 * (size_t)i + 2] = 0;
  F.solid[i] = 0;
}

// The workgroup's voxel sums: an LDS hash of voxel index -> integer channels
// (mass, three offsets, solid mass). Neighbouring slots are neighbouring
// particles, so a workgroup's 256 particles touch a few hundred voxels; summing
// them in LDS leaves one global atomic per voxel and channel instead of one per
// particle corner. A voxel that finds no free slot within a few probes adds
// straight to global memory. Wrap-around integer sums do not depend on order,
// so the result is the same either way.
constexpr uint32_t kSplatHash = 1024;
constexpr uint32_t kSplatProbes = 16;

struct SplatLds {
  uint32_t key[kSplatHash];  // voxel + 1, 0 = empty
  int ch[5][kSplatHash];
};

__device__ __forceinline__ void splat_add(SplatLds& L, const FieldDev& F, uint32_t c, bool fluid, const int v[4]) {
  uint32_t h = (c * 0x9E3779B1u) >> 22;
  for (uint32_t probe = 0; probe < kSplatProbes; ++probe) {
    const uint32_t prev = atomicCAS(&L.key[h], 0u, c + 1u);
    if (prev == 0u || prev == c + 1u) {
      if (fluid) {
        atomicAdd(&L.ch[0][h], v[0]);
        atomicAdd(&L.ch[1][h], v[1]);
        atomicAdd(&L.ch[2][h], v[2]);
        atomicAdd(&L.ch[3][h], v[3]);
      } else {
        atomicAdd(&L.ch[4][h], v[0]);
      }
      return;
    }
    h = (h + 1u) & (kSplatHash - 1u);
  }
  if (fluid) {
    atomicAdd(F.mass + c, v[0]);
    atomicAdd(F.off + 3 * (size_t)c + 0, v[1]);
    atomicAdd(F.off + 3 * (size_t)c + 1, v[2]);
    atomicAdd(F.off + 3 * (size_t)c + 2, v[3]);
  } else {
    atomicAdd(F.solid + c, v[0]);
  }
}

// splat (SurfaceField.ts:174-214), one thread per particle slot.
__global__ __launch_bounds__(PF_BLOCK) void k_field_splat(const SceneDev* __restrict__ sc, FieldDev F) {
  __shared__ SplatLds L;
  for (uint32_t i = threadIdx.x; i < kSplatHash; i += PF_BLOCK) {
    L.key[i] = 0u;
#pragma unroll
    for (int q = 0; q < 5; ++q) L.ch[q][i] = 0;
  }
  __syncthreads();
  const SceneDev& S = *sc;
  const uint32_t s = blockIdx.x * PF_BLOCK + threadIdx.x;
  const uint32_t id = s < S.n ? S.id[0][s] : 0xffffffffu;
  const bool fluid = s < S.n && fd_in(id, F.fluid_start, F.fluid_count);
  const bool solid = s < S.n && !fluid && fd_in(id, F.solid_start, F.solid_count);
  if (fluid || solid) {
    const float3 pos = pf_xyz(S.x[0][s]);
    const uint32_t samples = fluid ? F.samples : 1u;
    float3 trail = make_float3(0.f, 0.f, 0.f);
    if (samples > 1) {
      const float3 v = pf_xyz(S.v[0][s]);
      const float3 vv = make_float3(v.x * F.stretch, v.y * F.stretch, v.z * F.stretch);
      const float k = F.reach3 / fmaxf(pf_len3(vv), F.reach3);
      trail = make_float3(vv.x * k, vv.y * k, vv.z * k);
    }
    for (uint32_t sm = 0; sm < samples; ++sm) {
      const float frac = sm == 0 ? 0.0f : (float)sm / (float)(samples - 1);
      const float3 at =
          sm == 0 ? pos : make_float3(pos.x - trail.x * frac, pos.y - trail.y * frac, pos.z - trail.z * frac);
      const float g[3] = {(at.x - F.origin[0]) * F.inv_voxel - 0.5f, (at.y - F.origin[1]) * F.inv_voxel - 0.5f,
                          (at.z - F.origin[2]) * F.inv_voxel - 0.5f};
      bool inside = true;
#pragma unroll
      for (int a = 0; a < 3; ++a) inside = inside && g[a] >= 0.0f && g[a] < F.dims[a] - 1.0f;
      if (!inside) continue;
      float b[3], f[3];
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        b[a] = floorf(g[a]);
        f[a] = g[a] - b[a];
      }
#pragma unroll
      for (int z = 0; z < 2; ++z)
#pragma unroll
        for (int y = 0; y < 2; ++y)
#pragma unroll
          for (int x = 0; x < 2; ++x) {
            const float w = (((x ? f[0] : 1.0f - f[0]) * (y ? f[1] : 1.0f - f[1])) * (z ? f[2] : 1.0f - f[2])) *
                            (fluid ? F.inv_samples : 1.0f);
            // indexOf over float cell coordinates (SurfaceField.ts:147-152)
            const uint32_t c =
                (uint32_t)(((b[2] + (float)z) * F.nxny + (b[1] + (float)y) * F.nx) + (b[0] + (float)x));
            int v[4] = {pf_f2i_sat(rintf(w * kFixed)), 0, 0, 0};
            if (fluid) {
              v[1] = pf_f2i_sat(rintf((w * (f[0] - (float)x)) * kFixed));
              v[2] = pf_f2i_sat(rintf((w * (f[1] - (float)y)) * kFixed));
              v[3] = pf_f2i_sat(rintf((w * (f[2] - (float)z)) * kFixed));
            }
            splat_add(L, F, c, fluid, v);
          }
    }
  }
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < kSplatHash; i += PF_BLOCK) {
    const uint32_t key = L.key[i];
    if (!key) continue;
    const uint32_t c = key - 1u;
    if (L.ch[0][i]) atomicAdd(F.mass + c, L.ch[0][i]);
    if (L.ch[1][i]) atomicAdd(F.off + 3 * (size_t)c + 0, L.ch[1][i]);
    if (L.ch[2][i]) atomicAdd(F.off + 3 * (size_t)c + 1, L.ch[2][i]);
    if (L.ch[3][i]) atomicAdd(F.off + 3 * (size_t)c + 2, L.ch[3][i]);
    if (L.ch[4][i]) atomicAdd(F.solid + c, L.ch[4][i]);
  }
}

// wallKernel (SurfaceField.ts:251-285): the scene's primitives and SDF
// colliders as the wetted walls.
__global__ __launch_bounds__(PF_BLOCK) void k_field_walls(const SceneDev* __restrict__ sc, FieldDev F) {
  const SceneDev& S = *sc;
  const uint32_t i = blockIdx.x * PF_BLOCK + threadIdx.x;
  if (i >= F.count) return;
  float prox = 0.0f, nearest = 1e3f;
  if (F.wet && (S.nprim || S.nsdf)) {
    const uint32_t c[3] = {i % F.n[0], (i / F.n[0]) % F.n[1], i / (F.n[0] * F.n[1])};
    const float3 p = make_float3(F.origin[0] + ((float)c[0] + 0.5f) * F.voxel, F.origin[1] + ((float)c[1] + 0.5f) * F.voxel,
                                 F.origin[2] + ((float)c[2] + 0.5f) * F.voxel);
    for (uint32_t k = 0; k < S.nprim; ++k) nearest = fminf(nearest, prim_phi_now(S, k, p));
    for (uint32_t k = 0; k < S.nsdf; ++k) {
      const SdfDev& D = S.sdf[k];
      float phi;
      float3 g;
      sdf_sample(D, p, &phi, &g);
      // Outside its grid the field clamps; the distance to the grid's
      // bounding sphere is a valid lower bound there.
      float half[3], bo[3], ctr[3];
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        half[a] = (D.res[a] * D.voxel[a]) * 0.5f;
        bo[a] = D.origin[a] + half[a];
      }
#pragma unroll
      for (int r = 0; r < 3; ++r)
        ctr[r] = D.pos[r] + ((D.rot[3 * r] * bo[0] + D.rot[3 * r + 1] * bo[1]) + D.rot[3 * r + 2] * bo[2]) * D.scale;
      const float3 d = make_float3(p.x - ctr[0], p.y - ctr[1], p.z - ctr[2]);
      const float3 h3 = make_float3(half[0], half[1], half[2]);
      const float bound = pf_len3(d) - pf_len3(h3) * D.scale;
      nearest = fminf(nearest, fmaxf(phi, bound));
    }
    const float width = fmaxf(F.men_w, 1e-5f);
    prox = expf(-(fmaxf(nearest, 0.0f) / width));
  }
  F.walls[i] = make_float2(prox, nearest);
}

// blur (SurfaceField.ts:286-353) along AXIS; the last axis writes the field.
template <int AXIS>
__global__ __launch_bounds__(PF_BLOCK) void k_field_blur(FieldDev F) {
  const uint32_t i = blockIdx.x * PF_BLOCK + threadIdx.x;
  if (i >= F.count) return;
  const uint32_t nx = F.n[0], ny = F.n[1];
  const int c[3] = {(int)(i % nx), (int)((i / nx) % ny), (int)(i / (nx * ny))};
  float4 sum = make_float4(0.f, 0.f, 0.f, 0.f);
  float ssum = 0.0f;
  const int taps = (int)F.taps;
  for (int k = -taps; k <= taps; ++k) {
    int nb[3] = {c[0], c[1], c[2]};
    nb[AXIS] += k;
    if (nb[AXIS] < 0 || nb[AXIS] >= (int)F.n[AXIS]) continue;
    const uint32_t j = (uint32_t)nb[2] * (nx * ny) + (uint32_t)nb[1] * nx + (uint32_t)nb[0];
    float4 sm;
    float ss;
    if (AXIS == 0) {
      sm = make_float4((float)F.mass[j] / kFixed, (float)F.off[3 * (size_t)j + 0] / kFixed,
                       (float)F.off[3 * (size_t)j + 1] / kFixed, (float)F.off[3 * (size_t)j + 2] / kFixed);
      ss = (float)F.solid[j] / kFixed;
    } else {
      sm = (AXIS == 1 ? F.a : F.b)[j];
      ss = (AXIS == 1 ? F.sa : F.sb)[j];
    }
    // Re-base the neighbour's offsets onto this voxel's centre.
    const float sh = (float)k;
    const float4 reb = make_float4(sm.x, sm.y + (AXIS == 0 ? sh : 0.0f) * sm.x, sm.z + (AXIS == 1 ? sh : 0.0f) * sm.x,
                                   sm.w + (AXIS == 2 ? sh : 0.0f) * sm.x);
    const float wk = F.weights[k + taps];
    sum = make_float4(sum.x + reb.x * wk, sum.y + reb.y * wk, sum.z + reb.z * wk, sum.w + reb.w * wk);
    ssum = ssum + ss * wk;
  }
  if (AXIS < 2) {
    (AXIS == 0 ? F.a : F.b)[i] = sum;
    (AXIS == 0 ? F.sa : F.sb)[i] = ssum;
    return;
  }
  const float w = sum.x;
  const float den = fmaxf(w, 1e-8f);
  const float3 mean = make_float3(sum.y / den, sum.z / den, sum.w / den);
  const float radius = F.radius * F.inv_voxel;
  float phi = w > 1e-5f ? pf_len3(mean) - radius : kBand;
  phi = fminf(phi, kBand);
  const float2 wall = F.walls[i];
  const float prox = fmaxf(fminf(fmaxf(ssum * F.solid_scale, 0.0f), 1.0f), wall.x);
  // Wetting: lower phi near walls, only within reach of the surface.
  const float lift = F.men_h * F.inv_voxel;
  const float reach = fminf(fmaxf(1.0f - phi / fmaxf(lift * 1.5f, 1e-5f), 0.0f), 1.0f);
  phi = phi - (lift * prox) * (reach * reach);
  // No liquid inside solids.
  phi = fmaxf(phi, -wall.y * F.inv_voxel);
  F.out[i] = make_float4(phi, fminf(w * F.vol_scale, 4.0f), prox, 1.0f);
}

}  // namespace

struct pf_surface_field_s {
  pf_fluid_s* fluid = nullptr;
  pf_particles_s* ps = nullptr;
  FieldDev dev{};
  double voxel = 0, origin[3] = {0, 0, 0};
  bool disposed = false;
  // Frees whatever was allocated (also when creation fails part way).
  void release() {
    FieldDev& D = dev;
    for (void* b : {(void*)D.mass, (void*)D.off, (void*)D.solid, (void*)D.a, (void*)D.b, (void*)D.sa, (void*)D.sb,
                    (void*)D.walls, (void*)D.out})
      dfree(b);
    D.mass = D.off = D.solid = nullptr;
    D.a = D.b = D.out = nullptr;
    D.sa = D.sb = nullptr;
    D.walls = nullptr;
  }
  ~pf_surface_field_s() { release(); }
};

#define PF_API_BEGIN try {
#define PF_API_END                                    \
  }                                                   \
  catch (const pf::PfError& e) { return e.code; }     \
  catch (const pf::HipError& e) { return e.code; }    \
  catch (const std::bad_alloc&) {                     \
    pf::set_error("host allocation failed");          \
    return PF_ENOMEM;                                 \
  }                                                   \
  catch (const std::exception& e) {                   \
    pf::set_error("%s", e.what());                    \
    return PF_EINVAL;                                 \
  }                                                   \
  return PF_OK;

static void field_alive(pf_surface_field* f) {
  if (!f || f->disposed) fail(PF_ESTATE, "FluidSurfaceRenderer: the renderer has been disposed");
  if (f->fluid->disposed || f->ps->disposed) fail(PF_ESTATE, "FluidSurfaceRenderer: the fluid has been disposed");
  // A served scene's buffers belong to the server (and may be evicted).
  if (f->ps->server) fail(PF_ESTATE, "FluidSurfaceRenderer: the scene is being served; remove it from the server first");
}

extern "C" {

int pf_surface_field_create(pf_fluid* fluid, const pf_surface_field_options* o, pf_surface_field** out) {
  PF_API_BEGIN
  if (!fluid || fluid->disposed || fluid->ps->disposed)
    fail(PF_EINVAL, "FluidSurfaceRenderer: the fluid has been disposed");
  if (fluid->ps->server)
    fail(PF_ESTATE, "FluidSurfaceRenderer: the scene is being served; remove it from the server first");
  if (!o || !out) fail(PF_EINVAL, "FluidSurfaceRenderer: null argument");
  const double stretch = o->motion_stretch;
  if (!(stretch >= 0) || !std::isfinite(stretch))
    fail(PF_EINVAL, "FluidSurfaceRenderer: motionStretch must be ≥ 0, got %g", stretch);
  // defaultVoxelBudget (FluidSurfaceRenderer.ts:590-592)
  const double budget = o->voxel_budget > 0 ? o->voxel_budget
                        : fluid->count >= 50000 ? 1200000.0 : fluid->count >= 20000 ? 900000.0 : 600000.0;
  if (!(budget >= 1) || !std::isfinite(budget))
    fail(PF_EINVAL, "FluidSurfaceRenderer: voxelBudget must be ≥ 1, got %g", budget);
  const double sz[3] = {o->max[0] - o->min[0], o->max[1] - o->min[1], o->max[2] - o->min[2]};
  if (!(std::min(sz[0], std::min(sz[1], sz[2])) > 0)) fail(PF_EINVAL, "FluidSurfaceRenderer: bounds must have positive extent");
  pf_particles* ps = fluid->ps;
  if (o->solids_count && (o->solids_start > ps->n || o->solids_count > ps->n - o->solids_start))
    fail(PF_EINVAL, "FluidSurfaceRenderer: solids range [%u, +%u) is outside the particle system", o->solids_start,
         o->solids_count);
  auto F = std::make_unique<pf_surface_field_s>();
  F->fluid = fluid;
  F->ps = ps;
  // SurfaceField constructor (SurfaceField.ts:116-244), in double as the reference.
  const double r = ps->radius;
  const double spacing = fluid->spacing;
  const double voxel = std::max(r, std::cbrt((sz[0] * sz[1] * sz[2]) / budget));
  FieldDev& D = F->dev;
  for (int a = 0; a < 3; ++a) {
    D.n[a] = (uint32_t)std::ceil(sz[a] / voxel);
    D.origin[a] = (float)o->min[a];
    D.dims[a] = (float)D.n[a];
    F->origin[a] = o->min[a];
  }
  const uint64_t count = (uint64_t)D.n[0] * D.n[1] * D.n[2];
  if (count > (1ull << 24))
    fail(PF_EINVAL, "FluidSurfaceRenderer: %llu voxels; the grid's float cell index is exact to 2^24",
         (unsigned long long)count);
  D.count = (uint32_t)count;
  F->voxel = voxel;
  D.voxel = (float)voxel;
  D.inv_voxel = (float)(1.0 / voxel);
  D.nxny = (float)(D.n[0] * D.n[1]);
  D.nx = (float)D.n[0];
  const double sigma = (0.75 * spacing) / voxel;
  const int taps = (int)std::min(6.0, std::max(2.0, std::ceil(2.5 * sigma)));
  double wt[13], wsum = 0;
  for (int k = 0; k <= 2 * taps; ++k) {
    wt[k] = std::exp(-((double)(k - taps) * (double)(k - taps)) / (2 * sigma * sigma));
    wsum += wt[k];
  }
  for (int k = 0; k <= 2 * taps; ++k) D.weights[k] = (float)(wt[k] / wsum);
  D.taps = (uint32_t)taps;
  D.samples = stretch > 0 ? 3u : 1u;
  D.stretch = (float)stretch;
  D.reach3 = (float)(3 * spacing);
  D.inv_samples = (float)(1.0 / D.samples);
  const double vs = (spacing / voxel) * (spacing / voxel) * (spacing / voxel);
  D.vol_scale = (float)vs;
  D.solid_scale = (float)(vs * 1.5);
  D.fluid_start = fluid->start;
  D.fluid_count = fluid->count;
  D.solid_start = o->solids_start;
  D.solid_count = o->solids_count;
  D.wet = o->wet ? 1u : 0u;
  D.radius = (float)(o->radius >= 0 ? o->radius : r * 1.25);
  D.men_h = (float)(o->meniscus_height >= 0 ? o->meniscus_height : r * 1.6);
  D.men_w = (float)(o->meniscus_width >= 0 ? o->meniscus_width : r * 2.5);
  D.mass = dalloc<int>(count);
  D.off = dalloc<int>(3 * count);
  D.solid = dalloc<int>(count);
  D.a = dalloc<float4>(count);
  D.b = dalloc<float4>(count);
  D.sa = dalloc<float>(count);
  D.sb = dalloc<float>(count);
  D.walls = dalloc<float2>(count);
  D.out = dalloc<float4>(count);
  *out = F.release();
  PF_API_END
}

int pf_surface_field_info(pf_surface_field* f, uint32_t resolution[3], double origin[3], double* voxel) {
  PF_API_BEGIN
  field_alive(f);
  for (int a = 0; a < 3; ++a) {
    if (resolution) resolution[a] = f->dev.n[a];
    if (origin) origin[a] = f->origin[a];
  }
  if (voxel) *voxel = f->voxel;
  PF_API_END
}

int pf_surface_field_update(pf_surface_field* f) {
  PF_API_BEGIN
  field_alive(f);
  pf_particles* ps = f->ps;
  ps->upload_desc();
  hipStream_t st = ps->ctx->stream;
  const FieldDev& D = f->dev;
  const dim3 vox((D.count + PF_BLOCK - 1) / PF_BLOCK), parts((ps->n + PF_BLOCK - 1) / PF_BLOCK), blk(PF_BLOCK);
  Context* ctx = ps->ctx;
  {
    PF_LAUNCH(ctx, "field_clear");
    hipLaunchKernelGGL(k_field_clear, vox, blk, 0, st, D);
  }
  {
    PF_LAUNCH(ctx, "field_splat");
    hipLaunchKernelGGL(k_field_splat, parts, blk, 0, st, ps->cur(), D);
  }
  {
    PF_LAUNCH(ctx, "field_walls");
    hipLaunchKernelGGL(k_field_walls, vox, blk, 0, st, ps->cur(), D);
  }
  {
    PF_LAUNCH(ctx, "field_blur");
    hipLaunchKernelGGL(k_field_blur<0>, vox, blk, 0, st, D);
    hipLaunchKernelGGL(k_field_blur<1>, vox, blk, 0, st, D);
    hipLaunchKernelGGL(k_field_blur<2>, vox, blk, 0, st, D);
  }
  PF_HIP(hipGetLastError());
  PF_API_END
}

int pf_surface_field_readback(pf_surface_field* f, float* rgba) {
  PF_API_BEGIN
  field_alive(f);
  if (!rgba) fail(PF_EINVAL, "FluidSurfaceRenderer.readback: null output");
  PF_HIP(hipStreamSynchronize(f->ps->ctx->stream));
  PF_HIP(hipMemcpy(rgba, f->dev.out, (size_t)f->dev.count * sizeof(float4), hipMemcpyDeviceToHost));
  PF_API_END
}

int pf_surface_field_destroy(pf_surface_field* f) {
  PF_API_BEGIN
  if (!f || f->disposed) return PF_OK;
  (void)hipStreamSynchronize(f->ps->ctx->stream);
  // The handle stays valid (disposed) so a second destroy is harmless, as
  // for the library's other objects.
  f->release();
  f->disposed = true;
  PF_API_END
}

}  // extern "C"
