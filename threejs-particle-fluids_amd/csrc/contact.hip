// Particle-particle contacts (src/core/contact: generate.ts:39-88,
// solve.ts:33-189, ContactBuffer.ts:16-98; schedule loop.ts:346-418).
//
// Pairs are found once per substep, right after the sort, by a grid walk from
// every emitting particle; each record holds the sorted slots of its two
// particles. The per-contact passes (stabilise, solve, kinetic friction) run
// one thread per record in a grid-stride loop over the emitted count and
// scatter fixed-point corrections into a per-slot accumulator with integer
// atomics; the collider passes (k_walls, k_friction) start their own sums
// from it and apply both. Integer adds are exact and order independent, and
// each pass computes a pair's two corrections as exact negations of one
// another whichever end is stored first, so the result does not depend on
// the order the walk appended the pairs in.
#include "pf_internal.hpp"
#include "pf_kernels.hpp"
#include "collide.hpp"

namespace pf {

static __device__ __forceinline__ bool ct_in(uint32_t id, uint32_t start, uint32_t count) {
  return id >= start && id - start < count;
}

// A particle searches for contacts unless a material keeps its own particles
// apart (loop.ts:372-376): every fluid's range, plus the extra quiet ranges.
static __device__ __forceinline__ bool ct_emitter(const SceneDev& S, uint32_t id) {
  for (uint32_t f = 0; f < S.nfluids; ++f)
    if (ct_in(id, S.fluid[f].start, S.fluid[f].count)) return false;
  for (uint32_t q = 0; q < S.ct.nquiet; ++q)
    if (ct_in(id, S.ct.qstart[q], S.ct.qcount[q])) return false;
  return true;
}

static __device__ __forceinline__ int ct_lticks(float v) { return pf_f2i_sat(v * PF_CT_LAMBDA_SCALE); }

static __device__ __forceinline__ void ct_scatter(int4* acc, uint32_t s, float3 v, float scale) {
  int* a = reinterpret_cast<int*>(acc + s);
  atomicAdd(a + 0, pf_ticks(v.x, scale));
  atomicAdd(a + 1, pf_ticks(v.y, scale));
  atomicAdd(a + 2, pf_ticks(v.z, scale));
}

// ---------------------------------------------------------------------------
// Pair search (generate.ts:39-88). One thread per slot; emitters walk the
// buckets around their prediction with the reference's bucket dedup, skipping
// stencil cells that no particle within the candidate radius can occupy (the
// grid's cell is the widest query, often the fluid's smoothing radius; a wide
// margin covers the rounding of the cell assignment). A pair is stored once:
// by the emitter, by the lower original index when both emit. Candidates are
// tested by distance before their other attributes are read. The block
// appends its pairs to LDS and reserves their records with one global atomic;
// a thread whose pair does not fit in LDS appends it directly. The counter
// counts every pair, stored or not (ContactBuffer.ts:44-46).
// ---------------------------------------------------------------------------
#define PF_CT_LDS 1024

template <bool PERSIST>
__global__ __launch_bounds__(PF_BLOCK) void k_contact_generate(const SceneDev* __restrict__ scenes,
                                                               const Tile* __restrict__ tiles,
                                                               uint32_t ntiles, int p, int c) {
  __shared__ uint2 buf[PF_CT_LDS];
  __shared__ uint32_t nbuf, base;
  pf_for_tiles<PERSIST>(ntiles, [&](uint32_t ti) {
    const Tile t = tiles[ti];
    const SceneDev& S = scenes[t.scene];
    if (!S.ct.on) return;  // uniform over the block
    const uint32_t tid = threadIdx.x;
    if (tid == 0) nbuf = 0u;
    __syncthreads();
    const uint32_t s = t.first + tid;
    if (s < S.n) {
      const uint32_t I = S.id[p][s];
      if (ct_emitter(S, I)) {
        const float4* __restrict__ xs = S.xs[c];
        const float3 xi = pf_xyz(xs[s]);
        const uint32_t gi = S.grp[p][s];
        const float wi = S.w[p][s];
        const float r3[3] = {(xi.x - S.ox) / S.cell, (xi.y - S.oy) / S.cell, (xi.z - S.oz) / S.cell};
        int q[3];
        float lo[3], hi[3];
  #pragma unroll
        for (int a = 0; a < 3; ++a) {
          const float fl = floorf(r3[a]);
          q[a] = pf_f2i_sat(fl);
          lo[a] = (r3[a] - fl) * S.cell;
          hi[a] = (1.0f - (r3[a] - fl)) * S.cell;
        }
        const float lim = S.ct.cr + 1e-4f * S.cell;
        const float lim2 = lim * lim;
        uint32_t visited[27];
        uint32_t nv = 0;
        for (int dz = -1; dz <= 1; ++dz) {
          const float ez = dz < 0 ? lo[2] : (dz > 0 ? hi[2] : 0.0f);
          for (int dy = -1; dy <= 1; ++dy) {
            const float ey = dy < 0 ? lo[1] : (dy > 0 ? hi[1] : 0.0f);
            for (int dx = -1; dx <= 1; ++dx) {
              const float ex = dx < 0 ? lo[0] : (dx > 0 ? hi[0] : 0.0f);
              if ((ex * ex + ey * ey) + ez * ez > lim2) continue;
              const uint32_t b = pf_morton((int)((uint32_t)q[0] + (uint32_t)dx), (int)((uint32_t)q[1] + (uint32_t)dy),
                                           (int)((uint32_t)q[2] + (uint32_t)dz)) & S.hash_mask;
              bool seen = false;
              for (uint32_t k = 0; k < nv; ++k) seen |= visited[k] == b;
              visited[nv++] = b;
              if (seen) continue;
              const uint32_t en = S.start[b + 1];
              for (uint32_t j = S.start[b]; j < en; ++j) {
                const float3 xj = pf_xyz(xs[j]);
                const float3 d = make_float3(xi.x - xj.x, xi.y - xj.y, xi.z - xj.z);
                if (pf_dot3(d, d) >= S.ct.cr2) continue;
                const uint32_t J = S.id[p][j];
                if (J == I) continue;
                if (J < I && ct_emitter(S, J)) continue;
                if (gi != 0u && gi == S.grp[p][j]) continue;
                if (wi <= 0.0f && S.w[p][j] <= 0.0f) continue;
                const uint32_t k = atomicAdd(&nbuf, 1u);
                if (k < PF_CT_LDS) {
                  buf[k] = make_uint2(s, j);
                } else {
                  const uint32_t g = atomicAdd(S.ct.count, 1u);
                  if (g < S.ct.max) {
                    S.ct.pair[g] = make_uint2(s, j);
                    S.ct.lam[g] = make_int2(0, 0);
                  }
                }
              }
            }
          }
        }
      }
    }
    __syncthreads();
    const uint32_t nb = min(nbuf, (uint32_t)PF_CT_LDS);
    if (tid == 0 && nb) base = atomicAdd(S.ct.count, nb);
    __syncthreads();
    for (uint32_t k = tid; k < nb; k += PF_BLOCK) {
      const uint32_t g = base + k;
      if (g < S.ct.max) {
        S.ct.pair[g] = buf[k];
        // resetLambdaKernel (ContactBuffer.ts:67-71), for the records in use
        S.ct.lam[g] = make_int2(0, 0);
      }
    }
  });
}

// Records of this substep: contacts [t.first + tid, used) in steps of the
// scene's contact workgroups.
#define PF_CT_LOOP(S, t, cidx)                                                           \
  const uint32_t ct_used = min(*(S).ct.count, (S).ct.max);                               \
  const uint32_t ct_stride = (S).ct.blocks * PF_BLOCK;                                   \
  for (uint32_t cidx = (t).first + threadIdx.x; cidx < ct_used; cidx += ct_stride)

// ---------------------------------------------------------------------------
// Stabilisation (solve.ts:134-181): separate pairs that overlap at the start
// of the substep, moving committed and predicted positions alike.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(PF_BLOCK) void k_contact_stabilize(const SceneDev* __restrict__ scenes,
                                                                const Tile* __restrict__ tiles, int p) {
  const Tile t = tiles[blockIdx.x];
  const SceneDev& S = scenes[t.scene];
  PF_CT_LOOP(S, t, ci) {
    const uint2 pr = S.ct.pair[ci];
    const float3 xi = pf_xyz(S.x[p][pr.x]);
    const float3 xj = pf_xyz(S.x[p][pr.y]);
    const float3 off = make_float3(xi.x - xj.x, xi.y - xj.y, xi.z - xj.z);
    const float dist = pf_len3(off);
    const float wi = S.w[p][pr.x], wj = S.w[p][pr.y];
    const float ws = wi + wj;
    if (dist <= 1e-8f || dist >= S.ct.dist || ws <= 0.0f) continue;
    const float3 n = make_float3(off.x / dist, off.y / dist, off.z / dist);
    const float dl = (S.ct.dist - dist) / ws;
    const float ai = wi * dl, aj = wj * dl;
    ct_scatter(S.ct.acc, pr.x, make_float3(n.x * ai, n.y * ai, n.z * ai), S.pos_scale);
    ct_scatter(S.ct.acc, pr.y, make_float3(-(n.x * aj), -(n.y * aj), -(n.z * aj)), S.pos_scale);
  }
}

// The positions accumulator applied to the committed and the predicted
// positions (loop.ts:389-393). The predictions move to the other buffer, so
// the sorted copy the grid was built from stays behind for the fluid
// neighbour build (neighborList.ts:59). Scenes without contacts in the same
// launch only copy.
template <bool PERSIST>
__global__ __launch_bounds__(PF_BLOCK) void k_contact_stabilize_apply(const SceneDev* __restrict__ scenes,
                                                                      const Tile* __restrict__ tiles,
                                                                      uint32_t ntiles, int p, int c) {
  pf_for_tiles<PERSIST>(ntiles, [&](uint32_t ti) {
    const Tile t = tiles[ti];
    const SceneDev& S = scenes[t.scene];
    const uint32_t s = t.first + threadIdx.x;
    if (s >= S.n) return;
    float4 xs = S.xs[c][s];
    if (S.ct.on) {
      const int4 a = S.ct.acc[s];
      S.ct.acc[s] = make_int4(0, 0, 0, 0);
      const int tk[3] = {a.x, a.y, a.z};
      bool sat;
      const float3 ch = ticks_to_change(tk, S.pos_inv, &sat);
      if (sat) pf_flag_raise(S.flags + PF_FLAG_POSITIONS);
      float4 x = S.x[p][s];
      x.x = x.x + ch.x;
      x.y = x.y + ch.y;
      x.z = x.z + ch.z;
      S.x[p][s] = x;
      xs.x = xs.x + ch.x;
      xs.y = xs.y + ch.y;
      xs.z = xs.z + ch.z;
    }
    S.xs[c ^ 1][s] = xs;
  });
}

// ---------------------------------------------------------------------------
// Solve (solve.ts:33-96): non-penetration plus static friction, once per
// iteration, into the positions accumulator.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(PF_BLOCK) void k_contact_solve(const SceneDev* __restrict__ scenes,
                                                            const Tile* __restrict__ tiles, int p, int c) {
  const Tile t = tiles[blockIdx.x];
  const SceneDev& S = scenes[t.scene];
  PF_CT_LOOP(S, t, ci) {
    const uint2 pr = S.ct.pair[ci];
    const float3 pi = pf_xyz(S.xs[c][pr.x]);
    const float3 pj = pf_xyz(S.xs[c][pr.y]);
    const float3 off = make_float3(pi.x - pj.x, pi.y - pj.y, pi.z - pj.z);
    const float dist = pf_len3(off);
    const float wi = S.w[p][pr.x], wj = S.w[p][pr.y];
    const float ws = wi + wj;
    if (dist <= 1e-8f || dist >= S.ct.dist || ws <= 0.0f) continue;
    const float3 n = make_float3(off.x / dist, off.y / dist, off.z / dist);
    const float depth = S.ct.dist - dist;
    const float dln = depth / ws;
    int2 lam = S.ct.lam[ci];
    lam.x = pf_wrap_add(lam.x, ct_lticks(dln));
    S.ct.nrm[ci] = make_float4(n.x, n.y, n.z, 0.0f);
    const float3 ci3 = pf_xyz(S.x[p][pr.x]);
    const float3 cj3 = pf_xyz(S.x[p][pr.y]);
    const float3 slip = make_float3((pi.x - ci3.x) - (pj.x - cj3.x), (pi.y - ci3.y) - (pj.y - cj3.y),
                                    (pi.z - ci3.z) - (pj.z - cj3.z));
    const float3 tang = pf_sub3(slip, pf_mul3(n, pf_dot3(slip, n)));
    const float tl = pf_len3(tang);
    const float ln = (float)lam.x / PF_CT_LAMBDA_SCALE;
    const float lt = (float)lam.y / PF_CT_LAMBDA_SCALE;
    const float dlt = tl / ws;
    const bool sticks = tl > 1e-10f && dlt <= S.ct.mus * ln - lt;
    const float ai = wi * dln, aj = wj * dln;
    const float3 ni = make_float3(n.x * ai, n.y * ai, n.z * ai);
    const float3 nj = make_float3(-(n.x * aj), -(n.y * aj), -(n.z * aj));
    const float bi = wi / ws, bj = wj / ws;
    const float3 fi = make_float3(-(tang.x * bi), -(tang.y * bi), -(tang.z * bi));
    const float3 fj = make_float3(tang.x * bj, tang.y * bj, tang.z * bj);
    ct_scatter(S.ct.acc, pr.x, sticks ? pf_add3(ni, fi) : ni, S.pos_scale);
    ct_scatter(S.ct.acc, pr.y, sticks ? pf_add3(nj, fj) : nj, S.pos_scale);
    lam.y = pf_wrap_add(lam.y, ct_lticks(sticks ? dlt : 0.0f));
    S.ct.lam[ci] = lam;
  }
}

// ---------------------------------------------------------------------------
// Kinetic friction (solve.ts:147-189) into the velocities accumulator.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(PF_BLOCK) void k_contact_friction(const SceneDev* __restrict__ scenes,
                                                               const Tile* __restrict__ tiles, int p) {
  const Tile t = tiles[blockIdx.x];
  const SceneDev& S = scenes[t.scene];
  PF_CT_LOOP(S, t, ci) {
    const uint2 pr = S.ct.pair[ci];
    const float wi = S.w[p][pr.x], wj = S.w[p][pr.y];
    const float ws = wi + wj;
    const float ln = (float)S.ct.lam[ci].x / PF_CT_LAMBDA_SCALE;
    if (ws <= 0.0f || ln <= 0.0f) continue;
    const float3 n = pf_xyz(S.ct.nrm[ci]);
    const float3 vi = pf_xyz(S.v[p][pr.x]);
    const float3 vj = pf_xyz(S.v[p][pr.y]);
    const float3 rel = make_float3(vi.x - vj.x, vi.y - vj.y, vi.z - vj.z);
    const float3 vt = pf_sub3(rel, pf_mul3(n, pf_dot3(n, rel)));
    const float vl = pf_len3(vt);
    if (vl < 1e-6f) continue;
    const float change = fminf(((S.ct.muk * ln) * ws) / S.dt, vl);
    const float3 imp = make_float3((-((vt.x / vl) * change)) / ws, (-((vt.y / vl) * change)) / ws,
                                   (-((vt.z / vl) * change)) / ws);
    ct_scatter(S.ct.acc, pr.x, make_float3(imp.x * wi, imp.y * wi, imp.z * wi), S.vel_scale);
    ct_scatter(S.ct.acc, pr.y, make_float3(-(imp.x * wj), -(imp.y * wj), -(imp.z * wj)), S.vel_scale);
  }
}

// ---------------------------------------------------------------------------
// Host launchers
// ---------------------------------------------------------------------------
void launch_contact_generate(hipStream_t st, const SceneDev* sc, const Tile* tl, uint32_t ntiles, int p, int c) {
  if (!ntiles) return;
  launch_tiles(k_contact_generate<false>, k_contact_generate<true>, ntiles, 0, st, sc, tl, ntiles, p, c);
}

void launch_contact_stabilize(hipStream_t st, const SceneDev* sc, const Tile* tl, uint32_t ntiles,
                              const Tile* ctl, uint32_t nctiles, int p, int c) {
  if (nctiles) hipLaunchKernelGGL(k_contact_stabilize, dim3(nctiles), dim3(PF_BLOCK), 0, st, sc, ctl, p);
  launch_tiles(k_contact_stabilize_apply<false>, k_contact_stabilize_apply<true>, ntiles, 0, st, sc, tl, ntiles, p, c);
}

void launch_contact_solve(hipStream_t st, const SceneDev* sc, const Tile* ctl, uint32_t nctiles, int p, int c) {
  if (nctiles) hipLaunchKernelGGL(k_contact_solve, dim3(nctiles), dim3(PF_BLOCK), 0, st, sc, ctl, p, c);
}

void launch_contact_friction(hipStream_t st, const SceneDev* sc, const Tile* ctl, uint32_t nctiles, int p) {
  if (nctiles) hipLaunchKernelGGL(k_contact_friction, dim3(nctiles), dim3(PF_BLOCK), 0, st, sc, ctl, p);
}

}  // namespace pf
