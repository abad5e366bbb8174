// Position Based Fluids passes for gfx950, as cell-ordered gathers over the
// cell-sorted particle state (src/fluids/sim/*, FluidSystem.ts:200-309).
//
// Kernels (one thread per sorted slot; the per-substep neighbour list holds
// sorted slots, so neighbour reads land in nearby cache lines):
//   k_neighbors     27-bucket walk -> neighbour list, fused colour normal
//                   (neighborList.ts:45-72, cohesion.ts:25-39)
//   k_impulses      surface tension + adhesion as a gather with exact
//                   fixed-point sums, fused impulse apply
//                   (cohesion.ts:63-103, adhesion.ts:34-60, accumulator.ts:101-131)
//   k_lambda        density and lambda (lambda.ts:30-71)
//   k_delta         position correction + apply, optionally fused container
//                   contact + position accumulator apply and advect
//                   (positionDelta.ts:26-85, PrimitiveSet.ts:336-364, integrate.ts:36-44)
//   k_velocity      vorticity and XSPH walk (velocity.ts:36-70)
//   k_vorticity     vorticity gradient + velocity apply, optionally fused
//                   kinetic friction (velocity.ts:74-119, PrimitiveSet.ts:366-391)
// plus standalone k_walls / k_advect / k_friction for scenes that cannot fuse.
#include "collide.hpp"
#include "pf_internal.hpp"
#include "pf_kernels.hpp"

namespace pf {

static __device__ __forceinline__ bool in_range(uint32_t id, uint32_t start, uint32_t count) {
  return id - start < count;
}

// Neighbour mass psi_j (shared.ts:52-62).
static __device__ __forceinline__ float neighbor_mass(const SceneDev& S, const FluidDev& F, int p,
                                                      uint32_t j) {
  if (S.has_boundary) {
    const float vb = S.bvol[p][j];
    if (vb > 0.0f) return F.rho0 * vb;
  }
  return S.mass[j];
}

// Neighbour volume V_j (shared.ts:65-68).
static __device__ __forceinline__ float neighbor_volume(const SceneDev& S, const FluidDev& F, int p,
                                                        uint32_t j) {
  if (S.has_boundary) {
    const float vb = S.bvol[p][j];
    if (vb > 0.0f) return vb;
  }
  return F.pv;
}

static __device__ __forceinline__ void flag_set(const SceneDev& S, int parity, int f) {
  atomicOr(S.flags + parity * PF_NUM_FLAGS + f, 1u);
}

// ---------------------------------------------------------------------------
// Neighbour list (+ colour-field normal when surface tension is on)
// ---------------------------------------------------------------------------
template <bool NORMAL>
__global__ __launch_bounds__(PF_BLOCK) void k_neighbors(const SceneDev* __restrict__ scenes,
                                                        const Tile* __restrict__ tiles,
                                                        uint32_t ntiles, int fi, int p, int c) {
  const Tile t = tiles[pf_xcd_remap(blockIdx.x, ntiles)];
  const SceneDev& S = scenes[t.scene];
  const FluidDev& F = S.fluid[fi];
  const uint32_t s = t.first + threadIdx.x;
  if (s >= S.n) return;
  if (!in_range(S.id[p][s], F.start, F.count)) return;
  const float4* __restrict__ xs = S.xs[c];
  const float4 xi4 = xs[s];
  const float3 xi = pf_xyz(xi4);
  const float rx = (xi.x - S.ox) / S.cell, ry = (xi.y - S.oy) / S.cell, rz = (xi.z - S.oz) / S.cell;
  const int qx = pf_f2i_sat(floorf(rx));
  const int qy = pf_f2i_sat(floorf(ry));
  const int qz = pf_f2i_sat(floorf(rz));
  // Below 64 buckets two stencil cells can share a bucket: walk with the
  // reference's dedup and without culling.
  const bool small_table = S.H < 64u;
  // Distance from the particle to the faces of its own cell, to skip stencil
  // cells that no particle within h can occupy (a wide margin covers the
  // rounding of the cell assignment).
  const float lim = F.h + 1e-4f * S.cell;
  const float lim2 = lim * lim;
  float lo[3], hi[3];
  {
    const float f[3] = {rx - floorf(rx), ry - floorf(ry), rz - floorf(rz)};
    for (int a = 0; a < 3; ++a) {
      lo[a] = f[a] * S.cell;
      hi[a] = (1.0f - f[a]) * S.cell;
    }
  }
  uint32_t visited[27];
  uint32_t nvisited = 0;
  uint32_t count = 0;
  uint32_t last_key = 0xffffffffu;
  bool over = false;
  float3 nsum = make_float3(0.f, 0.f, 0.f);
  const uint32_t n = S.n;
  uint32_t* __restrict__ nl = F.nl;
  const bool want_normal = NORMAL && (F.effects & PF_FX_SURFACE_TENSION);
  for (int dz = -1; dz <= 1 && !over; ++dz) {
    const float ez = dz < 0 ? lo[2] : (dz > 0 ? hi[2] : 0.0f);
    for (int dy = -1; dy <= 1 && !over; ++dy) {
      const float ey = dy < 0 ? lo[1] : (dy > 0 ? hi[1] : 0.0f);
      for (int dx = -1; dx <= 1 && !over; ++dx) {
        const float ex = dx < 0 ? lo[0] : (dx > 0 ? hi[0] : 0.0f);
        const uint32_t b =
            pf_morton((int)((uint32_t)qx + (uint32_t)dx), (int)((uint32_t)qy + (uint32_t)dy),
                      (int)((uint32_t)qz + (uint32_t)dz)) &
            S.hash_mask;
        if (small_table) {
          bool seen = false;
          for (uint32_t k = 0; k < nvisited; ++k) seen |= (visited[k] == b);
          visited[nvisited++] = b;
          if (seen) continue;
        } else if ((ex * ex + ey * ey) + ez * ez > lim2) {
          continue;
        }
        const uint32_t stencil = (uint32_t)((dz + 1) * 9 + (dy + 1) * 3 + (dx + 1));
        const uint32_t st = S.start[b];
        const uint32_t en = S.start[b + 1];
        for (uint32_t j0 = st; j0 < en && !over; j0 += 4) {
          float4 xj[4];
#pragma unroll
          for (int u = 0; u < 4; ++u) xj[u] = (j0 + u < en) ? xs[j0 + u] : make_float4(3.0e38f, 0.f, 0.f, 0.f);
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            const uint32_t j = j0 + u;
            const float3 d = make_float3(xi.x - xj[u].x, xi.y - xj[u].y, xi.z - xj[u].z);
            const float r2 = pf_dot3(d, d);
            if (j < en && r2 < F.h2 && !over) {
              if (count < PF_MAX_NEIGHBORS) {
                nl[(size_t)count * n + s] = j;
                ++count;
                last_key = (stencil << 27) | j;
                if (want_normal) {
                  const float3 g = pf_spiky_grad(d, F.h, F.spiky);
                  nsum = pf_add3(nsum, pf_mul3(g, neighbor_volume(S, F, p, j)));
                }
              } else {
                // The list is full and another neighbour exists: it overflows,
                // nothing later in the walk can enter it.
                over = true;
              }
            }
          }
        }
      }
    }
  }
  F.nl_count[s] = count;
  // Walk position of the last kept neighbour of a truncated list; the
  // surface-tension pass uses it to know which partners this list dropped.
  F.nl_last[s] = (over && !small_table) ? last_key : 0xffffffffu;
  if (want_normal) F.normal[s] = make_float4(nsum.x * F.h, nsum.y * F.h, nsum.z * F.h, 0.0f);
  // Per-substep flags live in the half of the substep's pre-sort parity.
  if (over) flag_set(S, p ^ 1, PF_FLAG_NEIGHBORS);
}

// ---------------------------------------------------------------------------
// Surface tension (pairwise, gathered) + adhesion + impulse apply
// ---------------------------------------------------------------------------

// Cohesion spline C(r) (cohesion.ts:46-55).
static __device__ __forceinline__ float cohesion_spline(const FluidDev& F, float r) {
  const float gap = fmaxf(F.h - r, 0.0f);
  const float shared = ((gap * gap) * gap) * ((r * r) * r);
  const float value = (r * 2.0f > F.h) ? shared : (shared * 2.0f - F.coh_h6_64);
  return (r > 0.0f && r <= F.h) ? value * F.coh_coef : 0.0f;
}

// Velocity change of the pair (a, b) with id_a < id_b, as the lower particle's
// thread evaluates it (cohesion.ts:83-100). Returns false for coincident pairs.
static __device__ __forceinline__ bool pair_dv(const FluidDev& F, float dt, float3 xa, float3 xb,
                                               float3 na, float3 nb, float rhoa, float rhob,
                                               float3* dv) {
  const float3 off = pf_sub3(xa, xb);
  const float r = pf_len3(off);
  if (r < 1e-20f) return false;
  const float corr = (F.rho0 * 2.0f) / fmaxf(rhoa + rhob, 1.0f);
  const float cs = -(((F.gamma * F.mass) * F.mass) * cohesion_spline(F, r));
  const float3 coh = make_float3((off.x / r) * cs, (off.y / r) * cs, (off.z / r) * cs);
  const float gm = F.gamma * F.mass;
  const float3 curv = make_float3(-((na.x - nb.x) * gm), -((na.y - nb.y) * gm), -((na.z - nb.z) * gm));
  *dv = make_float3((((coh.x + curv.x) * corr) / F.mass) * dt, (((coh.y + curv.y) * corr) / F.mass) * dt,
                    (((coh.z + curv.z) * corr) / F.mass) * dt);
  return true;
}

// Adhesion spline A(r) (adhesion.ts:16-27).
static __device__ __forceinline__ float adhesion_spline(const FluidDev& F, float r) {
  const float root = sqrtf(sqrtf(fmaxf(((r * r) * F.adh_m4_over_h + r * 6.0f) - F.adh_2h, 0.0f)));
  return (r * 2.0f > F.h && r <= F.h) ? root * F.adh_coef : 0.0f;
}

template <bool ST, bool ADH>
__global__ __launch_bounds__(PF_BLOCK) void k_impulses(const SceneDev* __restrict__ scenes,
                                                       const Tile* __restrict__ tiles,
                                                       uint32_t ntiles, int fi, int p, int c) {
  const Tile t = tiles[pf_xcd_remap(blockIdx.x, ntiles)];
  const SceneDev& S = scenes[t.scene];
  const FluidDev& F = S.fluid[fi];
  const uint32_t s = t.first + threadIdx.x;
  if (s >= S.n) return;
  const float4* __restrict__ xs = S.xs[c];
  const float4 xi4 = xs[s];
  const uint32_t id = S.id[p][s];
  if (!in_range(id, F.start, F.count)) {
    S.xs[c ^ 1][s] = xi4;
    return;
  }
  const float3 xi = pf_xyz(xi4);
  const uint32_t n = S.n;
  const uint32_t cnt = F.nl_count[s];
  const uint32_t* __restrict__ nl = F.nl;
  int tk[3] = {0, 0, 0};
  float3 ni = make_float3(0.f, 0.f, 0.f);
  float rhoi = 0.0f;
  // A scene of a batch may lack an effect another scene has.
  const bool st_on = ST && (F.effects & PF_FX_SURFACE_TENSION);
  const bool adh_on = ADH && (F.effects & PF_FX_ADHESION);
  if (!(st_on || adh_on)) {
    S.xs[c ^ 1][s] = xi4;
    return;
  }
  if (st_on) {
    ni = pf_xyz(F.normal[s]);
    rhoi = S.rho[p][s];
  }
  const uint32_t fend = F.start + F.count;
  for (uint32_t k = 0; k < cnt; ++k) {
    const uint32_t j = nl[(size_t)k * n + s];
    if (st_on) {
      const uint32_t idj = S.id[p][j];
      if (idj > id && idj < fend) {
        float3 dv;
        if (pair_dv(F, S.dt, xi, pf_xyz(xs[j]), ni, pf_xyz(F.normal[j]), rhoi, S.rho[p][j], &dv)) {
          tk[0] = pf_wrap_add(tk[0], pf_ticks(dv.x, F.imp_scale));
          tk[1] = pf_wrap_add(tk[1], pf_ticks(dv.y, F.imp_scale));
          tk[2] = pf_wrap_add(tk[2], pf_ticks(dv.z, F.imp_scale));
        }
      } else if (idj < id && idj >= F.start) {
        // The pair belongs to j's thread in the reference; it exists there iff
        // this slot is in j's (possibly truncated) list. Lists keep their first
        // 64 neighbours in walk order, so compare this slot's walk key in j's
        // walk with the key of j's last kept neighbour.
        const uint32_t last = F.nl_last[j];
        bool listed = last == 0xffffffffu;
        if (!listed) {
          if (S.H < 64u) {
            for (uint32_t m = 0; m < PF_MAX_NEIGHBORS; ++m) listed |= (nl[(size_t)m * n + j] == s);
          } else {
            const float4 xj4 = xs[j];
            const int ddx = pf_query_coord(xi.x, S.ox, S.cell) - pf_query_coord(xj4.x, S.ox, S.cell);
            const int ddy = pf_query_coord(xi.y, S.oy, S.cell) - pf_query_coord(xj4.y, S.oy, S.cell);
            const int ddz = pf_query_coord(xi.z, S.oz, S.cell) - pf_query_coord(xj4.z, S.oz, S.cell);
            const uint32_t key = ((uint32_t)((ddz + 1) * 9 + (ddy + 1) * 3 + (ddx + 1)) << 27) | s;
            listed = key <= last;
          }
        }
        float3 dv;
        if (listed &&
            pair_dv(F, S.dt, pf_xyz(xs[j]), xi, pf_xyz(F.normal[j]), ni, S.rho[p][j], rhoi, &dv)) {
          tk[0] = pf_wrap_add(tk[0], pf_ticks(-dv.x, F.imp_scale));
          tk[1] = pf_wrap_add(tk[1], pf_ticks(-dv.y, F.imp_scale));
          tk[2] = pf_wrap_add(tk[2], pf_ticks(-dv.z, F.imp_scale));
        }
      }
    }
    if (adh_on) {
      const float vol = S.bvol[p][j];
      if (vol > 0.0f) {
        const float3 off = pf_sub3(xi, pf_xyz(xs[j]));
        const float r = pf_len3(off);
        if (!(r < 1e-20f)) {
          const float fs = -(((F.adhesion * F.mass) * (F.rho0 * vol)) * adhesion_spline(F, r));
          const float3 f = make_float3((off.x / r) * fs, (off.y / r) * fs, (off.z / r) * fs);
          tk[0] = pf_wrap_add(tk[0], pf_ticks((f.x / F.mass) * S.dt, F.imp_scale));
          tk[1] = pf_wrap_add(tk[1], pf_ticks((f.y / F.mass) * S.dt, F.imp_scale));
          tk[2] = pf_wrap_add(tk[2], pf_ticks((f.z / F.mass) * S.dt, F.imp_scale));
        }
      }
    }
  }
  bool sat;
  const float3 ch = ticks_to_change(tk, F.imp_inv, &sat);
  if (sat) flag_set(S, 0, PF_FLAG_IMPULSES);
  float4 v = S.v[p][s];
  v.x = v.x + ch.x;
  v.y = v.y + ch.y;
  v.z = v.z + ch.z;
  S.v[p][s] = v;
  S.xs[c ^ 1][s] = make_float4(xi.x + ch.x * S.dt, xi.y + ch.y * S.dt, xi.z + ch.z * S.dt, xi4.w);
}

// ---------------------------------------------------------------------------
// Density and lambda; lambda is written to xs[c].w for the correction pass
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(PF_BLOCK) void k_lambda(const SceneDev* __restrict__ scenes,
                                                     const Tile* __restrict__ tiles, uint32_t ntiles,
                                                     int fi, int p, int c) {
  const Tile t = tiles[pf_xcd_remap(blockIdx.x, ntiles)];
  const SceneDev& S = scenes[t.scene];
  const FluidDev& F = S.fluid[fi];
  const uint32_t s = t.first + threadIdx.x;
  if (s >= S.n) return;
  if (!in_range(S.id[p][s], F.start, F.count)) return;
  float4* __restrict__ xs = S.xs[c];
  const float3 xi = pf_xyz(xs[s]);
  const uint32_t n = S.n;
  const uint32_t cnt = F.nl_count[s];
  const uint32_t* __restrict__ nl = F.nl;
  float rho = 0.0f, gss = 0.0f;
  float3 gs = make_float3(0.f, 0.f, 0.f);
  // Four neighbours' loads in flight, then the sums in list order.
  for (uint32_t k0 = 0; k0 < cnt; k0 += 4) {
    uint32_t jj[4];
    float4 xj[4];
    float mj[4], wj[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) jj[u] = (k0 + u < cnt) ? nl[(size_t)(k0 + u) * n + s] : s;
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      xj[u] = xs[jj[u]];
      mj[u] = neighbor_mass(S, F, p, jj[u]);
      wj[u] = S.w[p][jj[u]];
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      if (k0 + u < cnt) {
        const float3 d = make_float3(xi.x - xj[u].x, xi.y - xj[u].y, xi.z - xj[u].z);
        rho = rho + mj[u] * pf_poly6(pf_dot3(d, d), F.h2, F.poly6);
        const float3 g = pf_mul3(pf_spiky_grad(d, F.h, F.spiky), mj[u]);
        gs = pf_add3(gs, g);
        gss = gss + pf_dot3(g, g) * wj[u];
      }
    }
  }
  S.rho[p][s] = rho;
  const float inv_rho0 = 1.0f / F.rho0;
  const float C = rho * inv_rho0 - 1.0f;
  float lambda = 0.0f;
  if (!(C <= 0.0f)) {
    const float gn = (pf_dot3(gs, gs) * S.w[p][s] + gss) * (inv_rho0 * inv_rho0);
    const float at = F.compliance / (S.dt * S.dt);
    lambda = -C / fmaxf(gn + at, 1e-20f);
  }
  xs[s].w = lambda;
}

// ---------------------------------------------------------------------------
// Position correction (+ container contact, accumulator apply, advect)
// ---------------------------------------------------------------------------
template <bool WALLS, bool ADVECT>
__global__ __launch_bounds__(PF_BLOCK) void k_delta(const SceneDev* __restrict__ scenes,
                                                    const Tile* __restrict__ tiles, uint32_t ntiles,
                                                    int fi, int p, int c, int first_iteration,
                                                    int substep) {
  const Tile t = tiles[pf_xcd_remap(blockIdx.x, ntiles)];
  const SceneDev& S = scenes[t.scene];
  const FluidDev& F = S.fluid[fi];
  const uint32_t s = t.first + threadIdx.x;
  if (s >= S.n) return;
  const float4* __restrict__ xs = S.xs[c];
  const float4 xi4 = xs[s];
  float3 xn = pf_xyz(xi4);
  const float w = S.w[p][s];
  if (in_range(S.id[p][s], F.start, F.count)) {
    const uint32_t n = S.n;
    const uint32_t cnt = F.nl_count[s];
    const uint32_t* __restrict__ nl = F.nl;
    const float li = xi4.w;
    float3 sum = make_float3(0.f, 0.f, 0.f);
    for (uint32_t k0 = 0; k0 < cnt; k0 += 4) {
      uint32_t jj[4];
      float4 xj[4];
      float mj[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) jj[u] = (k0 + u < cnt) ? nl[(size_t)(k0 + u) * n + s] : s;
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        xj[u] = xs[jj[u]];
        mj[u] = neighbor_mass(S, F, p, jj[u]);
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        if (k0 + u < cnt) {
          const float3 g = pf_spiky_grad(make_float3(xn.x - xj[u].x, xn.y - xj[u].y, xn.z - xj[u].z), F.h, F.spiky);
          const float sc = (li + xj[u].w) * mj[u];
          sum = pf_add3(sum, pf_mul3(g, sc));
        }
      }
    }
    xn = make_float3(xn.x + (sum.x / F.rho0) * w, xn.y + (sum.y / F.rho0) * w,
                     xn.z + (sum.z / F.rho0) * w);
  }
  if (WALLS || ADVECT) {
    float4 x = S.x[p][s];
    if (WALLS) {
      int tk[3] = {0, 0, 0};
      walls_solve(S, s, xn, pf_xyz(x), w, S.clock[substep], first_iteration != 0, tk);
      bool sat;
      const float3 ch = ticks_to_change(tk, S.pos_inv, &sat);
      if (sat) flag_set(S, 0, PF_FLAG_POSITIONS);
      xn = pf_add3(xn, ch);
    }
    if (ADVECT && w > 0.0f) {
      S.v[p][s] = make_float4((xn.x - x.x) / S.dt, (xn.y - x.y) / S.dt, (xn.z - x.z) / S.dt, 0.0f);
      S.x[p][s] = make_float4(xn.x, xn.y, xn.z, 0.0f);
    }
  }
  S.xs[c ^ 1][s] = make_float4(xn.x, xn.y, xn.z, 0.0f);
}

// Container contact + position accumulator apply (+ advect), in place.
template <bool ADVECT>
__global__ __launch_bounds__(PF_BLOCK) void k_walls(const SceneDev* __restrict__ scenes,
                                                    const Tile* __restrict__ tiles, uint32_t ntiles,
                                                    int p, int c, int first_iteration, int substep) {
  const Tile t = tiles[pf_xcd_remap(blockIdx.x, ntiles)];
  const SceneDev& S = scenes[t.scene];
  const uint32_t s = t.first + threadIdx.x;
  if (s >= S.n) return;
  const float4 xi4 = S.xs[c][s];
  float3 xn = pf_xyz(xi4);
  const float w = S.w[p][s];
  const float4 x = S.x[p][s];
  int tk[3] = {0, 0, 0};
  walls_solve(S, s, xn, pf_xyz(x), w, S.clock[substep], first_iteration != 0, tk);
  bool sat;
  const float3 ch = ticks_to_change(tk, S.pos_inv, &sat);
  if (sat) flag_set(S, 0, PF_FLAG_POSITIONS);
  xn = pf_add3(xn, ch);
  if (ADVECT && w > 0.0f) {
    S.v[p][s] = make_float4((xn.x - x.x) / S.dt, (xn.y - x.y) / S.dt, (xn.z - x.z) / S.dt, 0.0f);
    S.x[p][s] = make_float4(xn.x, xn.y, xn.z, 0.0f);
  }
  S.xs[c][s] = make_float4(xn.x, xn.y, xn.z, xi4.w);
}

// Advect only (integrate.ts:36-44), for scenes without colliders or fluid.
__global__ __launch_bounds__(PF_BLOCK) void k_advect(const SceneDev* __restrict__ scenes,
                                                     const Tile* __restrict__ tiles, uint32_t ntiles,
                                                     int p, int c) {
  const Tile t = tiles[pf_xcd_remap(blockIdx.x, ntiles)];
  const SceneDev& S = scenes[t.scene];
  const uint32_t s = t.first + threadIdx.x;
  if (s >= S.n) return;
  const float w = S.w[p][s];
  if (!(w > 0.0f)) return;
  const float4 xsv = S.xs[c][s];
  const float4 x = S.x[p][s];
  S.v[p][s] = make_float4((xsv.x - x.x) / S.dt, (xsv.y - x.y) / S.dt, (xsv.z - x.z) / S.dt, 0.0f);
  S.x[p][s] = make_float4(xsv.x, xsv.y, xsv.z, 0.0f);
}

// ---------------------------------------------------------------------------
// Velocity passes
// ---------------------------------------------------------------------------
template <bool VORT, bool VISC>
__global__ __launch_bounds__(PF_BLOCK) void k_velocity(const SceneDev* __restrict__ scenes,
                                                       const Tile* __restrict__ tiles,
                                                       uint32_t ntiles, int fi, int p) {
  const Tile t = tiles[pf_xcd_remap(blockIdx.x, ntiles)];
  const SceneDev& S = scenes[t.scene];
  const FluidDev& F = S.fluid[fi];
  const uint32_t s = t.first + threadIdx.x;
  if (s >= S.n) return;
  const bool vort = VORT && (F.effects & PF_FX_VORTICITY);
  const bool visc = VISC && (F.effects & PF_FX_VISCOSITY);
  if (!in_range(S.id[p][s], F.start, F.count)) {
    // |omega| of particles outside the fluid reads as zero in the gradient
    // pass (the reference's per-fluid buffer is zero there); slots change
    // particles every substep, so clear it.
    if (vort) F.omega[s] = make_float4(0.f, 0.f, 0.f, 0.f);
    return;
  }
  if (!(vort || visc)) return;
  const float4* __restrict__ X = S.x[p];
  const float4* __restrict__ V = S.v[p];
  const float3 xi = pf_xyz(X[s]);
  const float3 vi = pf_xyz(V[s]);
  const uint32_t n = S.n;
  const uint32_t cnt = F.nl_count[s];
  const uint32_t* __restrict__ nl = F.nl;
  float3 curl = make_float3(0.f, 0.f, 0.f);
  float3 sm = make_float3(0.f, 0.f, 0.f);
  for (uint32_t k0 = 0; k0 < cnt; k0 += 4) {
    uint32_t jj[4];
    float4 xj[4], vj4[4];
    float vol[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) jj[u] = (k0 + u < cnt) ? nl[(size_t)(k0 + u) * n + s] : s;
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      xj[u] = X[jj[u]];
      vj4[u] = V[jj[u]];
      vol[u] = neighbor_volume(S, F, p, jj[u]);
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      if (k0 + u < cnt) {
        const float3 vj = pf_xyz(vj4[u]);
        const float3 off = make_float3(xi.x - xj[u].x, xi.y - xj[u].y, xi.z - xj[u].z);
        if (vort) curl = pf_add3(curl, pf_mul3(pf_cross3(pf_sub3(vi, vj), pf_spiky_grad(off, F.h, F.spiky)), vol[u]));
        if (visc) sm = pf_add3(sm, pf_mul3(pf_mul3(pf_sub3(vj, vi), pf_poly6(pf_dot3(off, off), F.h2, F.poly6)), vol[u]));
      }
    }
  }
  if (vort) F.omega[s] = make_float4(curl.x, curl.y, curl.z, pf_len3(curl));
  if (visc) F.dv[s] = make_float4(sm.x * F.visc, sm.y * F.visc, sm.z * F.visc, 0.0f);
}

template <bool VORT, bool VISC, bool FRICTION>
__global__ __launch_bounds__(PF_BLOCK) void k_vorticity(const SceneDev* __restrict__ scenes,
                                                        const Tile* __restrict__ tiles,
                                                        uint32_t ntiles, int fi, int p, int c,
                                                        int substep, int touch_valid) {
  const Tile t = tiles[pf_xcd_remap(blockIdx.x, ntiles)];
  const SceneDev& S = scenes[t.scene];
  const FluidDev& F = S.fluid[fi];
  const uint32_t s = t.first + threadIdx.x;
  if (s >= S.n) return;
  float4 v = S.v[p][s];
  const bool vort = VORT && (F.effects & PF_FX_VORTICITY);
  const bool visc = VISC && (F.effects & PF_FX_VISCOSITY);
  if ((vort || visc) && in_range(S.id[p][s], F.start, F.count)) {
    float3 change = make_float3(0.f, 0.f, 0.f);
    if (vort) {
      const float4* __restrict__ X = S.x[p];
      const float3 xi = pf_xyz(X[s]);
      const float4 oi4 = F.omega[s];
      const float oi = oi4.w;
      const uint32_t n = S.n;
      const uint32_t cnt = F.nl_count[s];
      const uint32_t* __restrict__ nl = F.nl;
      float3 eta = make_float3(0.f, 0.f, 0.f);
      for (uint32_t k0 = 0; k0 < cnt; k0 += 4) {
        uint32_t jj[4];
        float4 xj[4];
        float oj[4];
        bool skip[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) jj[u] = (k0 + u < cnt) ? nl[(size_t)(k0 + u) * n + s] : s;
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          xj[u] = X[jj[u]];
          oj[u] = F.omega[jj[u]].w;
          skip[u] = (k0 + u >= cnt) || (S.has_boundary && S.bvol[p][jj[u]] > 0.0f);
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          if (!skip[u]) {
            const float3 g = pf_spiky_grad(make_float3(xi.x - xj[u].x, xi.y - xj[u].y, xi.z - xj[u].z), F.h, F.spiky);
            eta = pf_add3(eta, pf_mul3(g, (oj[u] - oi) * F.pv));
          }
        }
      }
      const float el = pf_len3(eta);
      if (el >= 1e-12f) {
        const float3 force = pf_cross3(make_float3(eta.x / el, eta.y / el, eta.z / el), pf_xyz(oi4));
        change = pf_add3(change, pf_mul3(pf_mul3(force, F.vort), S.dt));
      }
    }
    if (visc) change = pf_add3(change, pf_xyz(F.dv[s]));
    v.x = v.x + change.x;
    v.y = v.y + change.y;
    v.z = v.z + change.z;
  }
  if (FRICTION && touch_valid) {
    int tk[3] = {0, 0, 0};
    walls_friction(S, s, pf_xyz(S.xs[c][s]), pf_xyz(v), S.w[p][s], S.clock[substep], tk);
    bool sat;
    const float3 ch = ticks_to_change(tk, S.vel_inv, &sat);
    if (sat) flag_set(S, 0, PF_FLAG_VELOCITIES);
    v.x = v.x + ch.x;
    v.y = v.y + ch.y;
    v.z = v.z + ch.z;
  }
  S.v[p][s] = v;
}

// Friction + velocity accumulator apply on its own (no fluid to fuse with).
__global__ __launch_bounds__(PF_BLOCK) void k_friction(const SceneDev* __restrict__ scenes,
                                                       const Tile* __restrict__ tiles,
                                                       uint32_t ntiles, int p, int c, int substep) {
  const Tile t = tiles[pf_xcd_remap(blockIdx.x, ntiles)];
  const SceneDev& S = scenes[t.scene];
  const uint32_t s = t.first + threadIdx.x;
  if (s >= S.n) return;
  float4 v = S.v[p][s];
  int tk[3] = {0, 0, 0};
  walls_friction(S, s, pf_xyz(S.xs[c][s]), pf_xyz(v), S.w[p][s], S.clock[substep], tk);
  bool sat;
  const float3 ch = ticks_to_change(tk, S.vel_inv, &sat);
  if (sat) flag_set(S, 0, PF_FLAG_VELOCITIES);
  v.x = v.x + ch.x;
  v.y = v.y + ch.y;
  v.z = v.z + ch.z;
  S.v[p][s] = v;
}

// ---------------------------------------------------------------------------
// Host launchers
// ---------------------------------------------------------------------------
#define PF_GRID dim3(ntiles), dim3(PF_BLOCK), 0, st

void launch_neighbors(hipStream_t st, const SceneDev* sc, const Tile* tl, uint32_t ntiles, int fi,
                      int p, int c, bool normal) {
  if (!ntiles) return;
  if (normal) hipLaunchKernelGGL(k_neighbors<true>, PF_GRID, sc, tl, ntiles, fi, p, c);
  else hipLaunchKernelGGL(k_neighbors<false>, PF_GRID, sc, tl, ntiles, fi, p, c);
}

void launch_impulses(hipStream_t st, const SceneDev* sc, const Tile* tl, uint32_t ntiles, int fi,
                     int p, int c, bool surface_tension, bool adhesion) {
  if (!ntiles) return;
  if (surface_tension && adhesion) hipLaunchKernelGGL((k_impulses<true, true>), PF_GRID, sc, tl, ntiles, fi, p, c);
  else if (surface_tension) hipLaunchKernelGGL((k_impulses<true, false>), PF_GRID, sc, tl, ntiles, fi, p, c);
  else if (adhesion) hipLaunchKernelGGL((k_impulses<false, true>), PF_GRID, sc, tl, ntiles, fi, p, c);
  else hipLaunchKernelGGL((k_impulses<false, false>), PF_GRID, sc, tl, ntiles, fi, p, c);
}

void launch_lambda(hipStream_t st, const SceneDev* sc, const Tile* tl, uint32_t ntiles, int fi,
                   int p, int c) {
  if (!ntiles) return;
  hipLaunchKernelGGL(k_lambda, PF_GRID, sc, tl, ntiles, fi, p, c);
}

void launch_delta(hipStream_t st, const SceneDev* sc, const Tile* tl, uint32_t ntiles, int fi,
                  int p, int c, bool walls, bool advect, int first_iteration, int substep) {
  if (!ntiles) return;
  if (walls && advect) hipLaunchKernelGGL((k_delta<true, true>), PF_GRID, sc, tl, ntiles, fi, p, c, first_iteration, substep);
  else if (walls) hipLaunchKernelGGL((k_delta<true, false>), PF_GRID, sc, tl, ntiles, fi, p, c, first_iteration, substep);
  else if (advect) hipLaunchKernelGGL((k_delta<false, true>), PF_GRID, sc, tl, ntiles, fi, p, c, first_iteration, substep);
  else hipLaunchKernelGGL((k_delta<false, false>), PF_GRID, sc, tl, ntiles, fi, p, c, first_iteration, substep);
}

void launch_walls(hipStream_t st, const SceneDev* sc, const Tile* tl, uint32_t ntiles, int p, int c,
                  bool advect, int first_iteration, int substep) {
  if (!ntiles) return;
  if (advect) hipLaunchKernelGGL(k_walls<true>, PF_GRID, sc, tl, ntiles, p, c, first_iteration, substep);
  else hipLaunchKernelGGL(k_walls<false>, PF_GRID, sc, tl, ntiles, p, c, first_iteration, substep);
}

void launch_advect(hipStream_t st, const SceneDev* sc, const Tile* tl, uint32_t ntiles, int p, int c) {
  if (!ntiles) return;
  hipLaunchKernelGGL(k_advect, PF_GRID, sc, tl, ntiles, p, c);
}

void launch_velocity(hipStream_t st, const SceneDev* sc, const Tile* tl, uint32_t ntiles, int fi,
                     int p, bool vort, bool visc) {
  if (!ntiles || !(vort || visc)) return;
  if (vort && visc) hipLaunchKernelGGL((k_velocity<true, true>), PF_GRID, sc, tl, ntiles, fi, p);
  else if (vort) hipLaunchKernelGGL((k_velocity<true, false>), PF_GRID, sc, tl, ntiles, fi, p);
  else hipLaunchKernelGGL((k_velocity<false, true>), PF_GRID, sc, tl, ntiles, fi, p);
}

void launch_vorticity(hipStream_t st, const SceneDev* sc, const Tile* tl, uint32_t ntiles, int fi,
                      int p, int c, bool vort, bool visc, bool friction, int substep,
                      int touch_valid) {
  if (!ntiles) return;
#define PF_VK(a, b, f) hipLaunchKernelGGL((k_vorticity<a, b, f>), PF_GRID, sc, tl, ntiles, fi, p, c, substep, touch_valid)
  if (friction) {
    if (vort && visc) PF_VK(true, true, true);
    else if (vort) PF_VK(true, false, true);
    else if (visc) PF_VK(false, true, true);
    else PF_VK(false, false, true);
  } else {
    if (vort && visc) PF_VK(true, true, false);
    else if (vort) PF_VK(true, false, false);
    else if (visc) PF_VK(false, true, false);
  }
#undef PF_VK
}

void launch_friction(hipStream_t st, const SceneDev* sc, const Tile* tl, uint32_t ntiles, int p,
                     int c, int substep) {
  if (!ntiles) return;
  hipLaunchKernelGGL(k_friction, PF_GRID, sc, tl, ntiles, p, c, substep);
}

}  // namespace pf
