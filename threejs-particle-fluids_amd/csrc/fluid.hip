// Position Based Fluids passes for gfx950, as cell-ordered gathers over the
// cell-sorted particle state (src/fluids/sim/*, FluidSystem.ts:200-309).
//
// Kernels (one thread per sorted slot; the per-substep neighbour list holds
// sorted slots, so neighbour reads land in nearby cache lines):
//   k_neighbors     27-bucket walk -> neighbour list, fused colour normal
//                   (neighborList.ts:45-72, cohesion.ts:25-39)
//   k_impulses      surface tension, each pair from its lower end with the
//                   partner's half scattered (exact fixed-point sums), and
//                   adhesion; k_impulse_apply applies the sums
//                   (cohesion.ts:63-103, adhesion.ts:34-60, accumulator.ts:101-131)
//   k_lambda        density and lambda (lambda.ts:30-71)
//   k_delta         position correction + apply, optionally fused container
//                   contact + position accumulator apply and advect
//                   (positionDelta.ts:26-85, PrimitiveSet.ts:336-364, integrate.ts:36-44)
//   k_velocity      vorticity and XSPH walk (velocity.ts:36-70)
//   k_vorticity     vorticity gradient + velocity apply, optionally fused
//                   kinetic friction (velocity.ts:74-119, PrimitiveSet.ts:366-391)
// plus standalone k_walls / k_advect / k_friction for scenes that cannot fuse.
#include <type_traits>

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
  pf_flag_raise(S.flags + parity * PF_NUM_FLAGS + f);
}

// ---------------------------------------------------------------------------
// Block helpers
// ---------------------------------------------------------------------------
static __device__ __forceinline__ uint32_t wave_inclusive_scan(uint32_t v) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t y = __shfl_up(v, o, 64);
    if (lane >= o) v += y;
  }
  return v;
}

// Exclusive prefix of one value per thread over the block; *total gets the
// sum. sh holds PF_BLOCK / 64 words. Every thread of the block must call it.
static __device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t* sh,
                                                                uint32_t* total) {
  const uint32_t inc = wave_inclusive_scan(v);
  const int wave = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 63) sh[wave] = inc;
  __syncthreads();
  uint32_t before = 0, all = 0;
#pragma unroll
  for (int w = 0; w < PF_BLOCK / 64; ++w) {
    const uint32_t x = sh[w];
    if (w < wave) before += x;
    all += x;
  }
  *total = all;
  __syncthreads();
  return before + inc - v;
}

// Block-wide exclusive max of v (0 for thread 0); sh holds PF_BLOCK/64 words.
static __device__ __forceinline__ uint32_t block_exclusive_max(uint32_t v, uint32_t* sh) {
  const uint32_t lane = __lane_id();
  const int wave = threadIdx.x >> 6;
  uint32_t inc = v;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const uint32_t y = __shfl_up(inc, d);
    if (lane >= (uint32_t)d) inc = max(inc, y);
  }
  uint32_t exc = __shfl_up(inc, 1);
  if (lane == 0) exc = 0u;
  if (lane == 63) sh[wave] = inc;
  __syncthreads();
  uint32_t before = 0;
#pragma unroll
  for (int w = 0; w < PF_BLOCK / 64; ++w)
    if (w < wave) before = max(before, sh[w]);
  __syncthreads();
  return max(before, exc);
}

// Hash-set slot of a bucket (Fibonacci hashing into PF_TU_HASH slots).
static __device__ __forceinline__ uint32_t tu_hash(uint32_t b) {
  return (b * 0x9E3779B1u) >> (32 - 10);
}
static_assert(PF_TU_HASH == 1024, "tu_hash assumes 1024 slots");

static __device__ __forceinline__ float3 sel3(bool ok, float3 a, float3 b) {
  return make_float3(ok ? a.x : b.x, ok ? a.y : b.y, ok ? a.z : b.z);
}

// Colour-field normal n_i = h * sum_j V_j grad W (cohesion.ts:25-39) over a
// finished global-slot list, in list order.
static __device__ __forceinline__ float4 color_normal_global(const SceneDev& S, const FluidDev& F, int p,
                                                             const float4* __restrict__ xs, float3 xi,
                                                             uint32_t s, uint32_t count) {
  const uint32_t n = S.n;
  float3 nsum = make_float3(0.f, 0.f, 0.f);
  for (uint32_t k0 = 0; k0 < count; k0 += 4) {
    uint32_t jj[4];
    float4 xj[4];
    float vol[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) jj[u] = (k0 + u < count) ? F.nl[(size_t)(k0 + u) * n + s] : s;
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      xj[u] = xs[jj[u]];
      vol[u] = neighbor_volume(S, F, p, jj[u]);
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const float3 d = make_float3(xi.x - xj[u].x, xi.y - xj[u].y, xi.z - xj[u].z);
      nsum = sel3(k0 + u < count, pf_add3(nsum, pf_mul3(d, pf_spiky_scale(d, F.h, F.spiky) * vol[u])), nsum);
    }
  }
  return make_float4(nsum.x * F.h, nsum.y * F.h, nsum.z * F.h, 0.0f);
}

// Sorted slots of the tile's union entries threadIdx.x + i * PF_BLOCK, all
// loads in flight at once; entries past U read slot 0 (callers skip them).
#define PF_TU_PER (PF_TU_MAX / PF_BLOCK)
static __device__ __forceinline__ void tu_load_slots(const uint32_t* __restrict__ slots, uint32_t U,
                                                     uint32_t js[PF_TU_PER]) {
  const PF_G uint32_t* g = pf_g(slots);
#pragma unroll
  for (int i = 0; i < PF_TU_PER; ++i) {
    const uint32_t u = threadIdx.x + i * PF_BLOCK;
    js[i] = u < U ? g[u] : 0u;
  }
}

// Slot of key in an open-addressing table of hs entries (linear probing).
static __device__ __forceinline__ uint32_t cu_home(uint32_t key, uint32_t hs) {
  return __umulhi(key * 0x9E3779B1u, hs);
}

// Union of the particles named by a global-walk tile's finished lists (global
// slots, F.nl). Every thread of the block calls it; `count` is the calling
// thread's list length (0 for threads without a list). When at most `limit`
// distinct slots appear, they become the tile's union (F.tu_slots, slots
// ascending), each list is rewritten as 16-bit union indices (F.nl16) and
// the union size is returned; otherwise PF_TU_GLOBAL. The table lives in the
// union's LDS, 3 * cap words: 2 * cap keys (slot + 1) and as many 16-bit
// dense indices; `dense` (limit words of LDS) sorts the keys. Ascending
// slots make consecutive union entries mostly consecutive slots, so the
// passes that stage the union read it coalesced.
static __device__ uint32_t compact_union(const SceneDev& S, const FluidDev& F, uint32_t tile, uint32_t* lds,
                                         uint32_t cap, uint32_t limit, uint32_t* ctl, uint32_t count,
                                         uint32_t s, uint32_t* dense) {
  const uint32_t tid = threadIdx.x;
  const uint32_t hs = 2u * cap;
  uint32_t* hk = lds;                                        // [hs] slot + 1, 0 = empty
  uint16_t* hv = reinterpret_cast<uint16_t*>(lds + hs);      // [hs] union index
  const uint32_t n = S.n;
  __syncthreads();
  for (uint32_t i = tid; i < hs; i += PF_BLOCK) hk[i] = 0u;
  if (tid < 2) ctl[tid] = 0u;  // distinct slots, give-up flag
  __syncthreads();
  const PF_G uint32_t* __restrict__ col = pf_g(F.nl) + s;
  {
    // Eight list entries loaded per round trip; their home slots are claimed
    // together and only a collision probes on (a list holds distinct slots).
    volatile uint32_t* give_up = &ctl[1];
    for (uint32_t k0 = 0; k0 < count && !*give_up; k0 += 8u) {
      uint32_t keys[8], h[8], prev[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) keys[u] = col[(size_t)min(k0 + u, count - 1u) * n] + 1u;
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        h[u] = cu_home(keys[u], hs);
        prev[u] = k0 + u < count ? atomicCAS(&hk[h[u]], 0u, keys[u]) : keys[u];
      }
      uint32_t added = 0;
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        uint32_t pv = prev[u], hh = h[u];
        for (uint32_t probe = 1; probe < hs && pv != 0u && pv != keys[u]; ++probe) {
          hh = (hh + 1u == hs) ? 0u : hh + 1u;
          pv = atomicCAS(&hk[hh], 0u, keys[u]);
        }
        added += pv == 0u ? 1u : 0u;
      }
      if (added && atomicAdd(&ctl[0], added) + added > limit) *give_up = 1u;
    }
  }
  __syncthreads();
  // Uniform over the block.
  if (ctl[1] != 0u || ctl[0] > limit || limit == 0u) return PF_TU_GLOBAL;
  // The keys, gathered in table order (each thread a contiguous chunk).
  const uint32_t per = (hs + PF_BLOCK - 1u) / PF_BLOCK;
  const uint32_t b0 = min(tid * per, hs), b1 = min(b0 + per, hs);
  uint32_t mine = 0;
  for (uint32_t i = b0; i < b1; ++i) mine += hk[i] ? 1u : 0u;
  uint32_t U = 0;
  uint32_t at = block_exclusive_scan(mine, ctl + 4, &U);
  for (uint32_t i = b0; i < b1; ++i) {
    const uint32_t key = hk[i];
    if (key) dense[at++] = key;
  }
  __syncthreads();
  // Sorted ascending by a bitonic network over the next power of two P, in
  // its all-ascending form (each merge opens with the mirrored comparison):
  // entries past U stand for +inf and never move, so only pairs inside U swap.
  uint32_t P = 1;
  while (P < U) P <<= 1;
  for (uint32_t k = 2; k <= P; k <<= 1) {
    for (uint32_t j = k >> 1; j > 0; j >>= 1) {
      for (uint32_t cidx = tid; cidx < P / 2; cidx += PF_BLOCK) {
        const uint32_t a = (cidx / j) * 2u * j + cidx % j;
        const uint32_t b = j == (k >> 1) ? a ^ (k - 1u) : a + j;
        if (b < U) {
          const uint32_t x = dense[a], y = dense[b];
          if (x > y) {
            dense[a] = y;
            dense[b] = x;
          }
        }
      }
      __syncthreads();
    }
  }
  // Each key's table entry takes its rank.
  PF_G uint32_t* __restrict__ slots = pf_g(F.tu_slots) + (size_t)tile * PF_TU_MAX;
  for (uint32_t i = tid; i < U; i += PF_BLOCK) {
    const uint32_t key = dense[i];
    uint32_t h = cu_home(key, hs);
    for (uint32_t probe = 1; probe < hs && hk[h] != key; ++probe) h = (h + 1u == hs) ? 0u : h + 1u;
    hv[h] = (uint16_t)i;
    slots[i] = key - 1u;
  }
  __syncthreads();
  // The lists as pairs of union indices, column-major like the staged walk's.
  // Eight entries loaded per round trip; entries past the count are zero.
  PF_G uint32_t* __restrict__ row = pf_g(F.nl16) + s;
  for (uint32_t k0 = 0; k0 < count; k0 += 8u) {
    uint32_t keys[8], h[8], v[8], idx[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) keys[u] = col[(size_t)min(k0 + u, count - 1u) * n] + 1u;
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      h[u] = cu_home(keys[u], hs);
      v[u] = hk[h[u]];
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      for (uint32_t probe = 1; probe < hs && v[u] != keys[u]; ++probe) {
        h[u] = (h[u] + 1u == hs) ? 0u : h[u] + 1u;
        v[u] = hk[h[u]];
      }
      idx[u] = k0 + u < count ? (uint32_t)hv[h[u]] : 0u;
    }
#pragma unroll
    for (int u = 0; u < 8; u += 2)
      if (k0 + u < count) row[(size_t)((k0 + u) >> 1) * n] = idx[u] | (idx[u + 1] << 16);
  }
  return U;
}

// Union of the particles named by a global-walk tile's finished lists (global
// slots, F.nl) when the tile's bucket layout survived the walk (h_key, h_off,
// h_start; ulay <= 2^16 entries): a listed slot j lies in bucket tmp_key[j],
// at layout index (its union offset) + (j - its first slot). Every thread of
// the block calls it; `count` is the thread's list length (0 without a
// list). The listed entries are marked in an LDS bitmap over the layout and
// numbered in layout order by a prefix of the words' popcounts, so the union
// is in bucket order, slots ascending inside each bucket, and the passes that
// stage it read it coalesced. lds holds the bitmap (nw words), a 16-bit
// prefix per word and PF_TU_MAX claim bits (compact_layout_words). When at
// most `limit` entries are listed they become the tile's union (F.tu_slots,
// each slot written by the first list entry to claim it) and each list is
// rewritten as 16-bit union indices (F.nl16); the union size is returned.
// Otherwise the lists keep their global slots and PF_TU_GLOBAL is returned.
static __device__ __forceinline__ uint32_t compact_layout_words(uint32_t ulay) {
  const uint32_t nw = (ulay + 31u) >> 5;
  return nw + (nw + 1u) / 2u + PF_TU_MAX / 32u;
}
static __device__ uint32_t compact_layout(const SceneDev& S, const FluidDev& F, uint32_t tile, const uint32_t* h_key,
                                          const uint32_t* h_off, const uint32_t* h_start, uint32_t ulay,
                                          uint32_t* lds, uint32_t limit, uint32_t* ctl, uint32_t count, uint32_t s) {
  const uint32_t tid = threadIdx.x;
  const uint32_t n = S.n;
  const uint32_t nw = (ulay + 31u) >> 5;
  uint32_t* bitmap = lds;
  uint16_t* pre16 = reinterpret_cast<uint16_t*>(lds + nw);
  uint32_t* claim = lds + nw + (nw + 1u) / 2u;
  const uint32_t ul = ulay ? ulay - 1u : 0u;
  __syncthreads();  // every wave has left the walk (its marks share these words)
  for (uint32_t w = tid; w < nw; w += PF_BLOCK) bitmap[w] = 0u;
  if (tid < PF_TU_MAX / 32) claim[tid] = 0u;
  __syncthreads();
  const PF_G uint32_t* __restrict__ col = pf_g(F.nl) + s;
  const PF_G uint32_t* __restrict__ g_key = pf_g(S.tmp_key);
  // Layout index of a listed slot.
  auto layout_index = [&](uint32_t j, uint32_t b) {
    uint32_t h = tu_hash(b);
    for (uint32_t probe = 0; probe < PF_TU_HASH && h_key[h] != b + 1u; ++probe) h = (h + 1u) & (PF_TU_HASH - 1u);
    return min((h_off[h] >> 16) + (j - h_start[h]), ul);  // the walk's buckets are all in the set
  };
  // Eight list entries per round trip, their buckets' loads in flight
  // together; the layout indices are kept in the thread's own F.nl16 column
  // (rows written across the wave's lanes at once) for the numbering pass.
  PF_G uint32_t* __restrict__ out16 = pf_g(F.nl16) + s;
  for (uint32_t k0 = 0; k0 < count; k0 += 8u) {
    uint32_t js[8], bs[8], us[8];
#pragma unroll
    for (uint32_t e = 0; e < 8; ++e) js[e] = col[(size_t)min(k0 + e, count - 1u) * n];
#pragma unroll
    for (uint32_t e = 0; e < 8; ++e) bs[e] = g_key[js[e]];
#pragma unroll
    for (uint32_t e = 0; e < 8; ++e) {
      us[e] = layout_index(js[e], bs[e]);
      if (k0 + e < count) atomicOr(bitmap + (us[e] >> 5), 1u << (us[e] & 31u));
    }
#pragma unroll
    for (uint32_t e = 0; e < 8; e += 2)
      if (k0 + e < count) out16[(size_t)((k0 + e) >> 1) * n] = us[e] | (us[e + 1] << 16);
  }
  __syncthreads();
  constexpr uint32_t WPT = 2048u / PF_BLOCK;  // bitmap words per thread (2^16 bits)
  const uint32_t w0 = tid * WPT;
  uint32_t pc[WPT], sum = 0;
#pragma unroll
  for (uint32_t i = 0; i < WPT; ++i) {
    pc[i] = w0 + i < nw ? (uint32_t)__popc(bitmap[w0 + i]) : 0u;
    sum += pc[i];
  }
  uint32_t U = 0;
  uint32_t at = block_exclusive_scan(sum, ctl + 4, &U);
  // Uniform over the block.
  if (U > limit || limit == 0u) return PF_TU_GLOBAL;
#pragma unroll
  for (uint32_t i = 0; i < WPT; ++i) {
    if (w0 + i < nw) pre16[w0 + i] = (uint16_t)at;
    at += pc[i];
  }
  __syncthreads();
  // Each entry's union index; the first list entry to claim a union entry
  // writes its slot.
  PF_G uint32_t* __restrict__ slots = pf_g(F.tu_slots) + (size_t)tile * PF_TU_MAX;
  constexpr uint32_t rmax = PF_MAX_NEIGHBORS / 2 - 1;
  for (uint32_t k0 = 0; k0 < count; k0 += 8u) {
    uint32_t js[8], w[4], idx[8];
#pragma unroll
    for (uint32_t e = 0; e < 8; ++e) js[e] = col[(size_t)min(k0 + e, count - 1u) * n];
#pragma unroll
    for (uint32_t i = 0; i < 4; ++i) w[i] = out16[(size_t)min((k0 >> 1) + i, rmax) * n];
#pragma unroll
    for (uint32_t e = 0; e < 8; ++e) {
      const uint32_t u = min((w[e >> 1] >> ((e & 1u) * 16u)) & 0xffffu, ul);
      const uint32_t below = bitmap[u >> 5] & ((1u << (u & 31u)) - 1u);
      idx[e] = k0 + e < count ? (uint32_t)pre16[u >> 5] + (uint32_t)__popc(below) : 0u;
      if (k0 + e < count && idx[e] < PF_TU_MAX) {
        const uint32_t bit = 1u << (idx[e] & 31u);
        if (!(atomicOr(claim + (idx[e] >> 5), bit) & bit)) slots[idx[e]] = js[e];
      }
    }
#pragma unroll
    for (uint32_t e = 0; e < 8; e += 2)
      if (k0 + e < count) out16[(size_t)((k0 + e) >> 1) * n] = idx[e] | (idx[e + 1] << 16);
  }
  return U;
}

// ---------------------------------------------------------------------------
// Run walk: the neighbour lists of a wave's particles, one cell run at a time.
//
// The wave's active slots split into runs of equal query cell, cut at the
// wave's edges (so a run holds at most 64 particles). For a run the wave
// enumerates the candidates of the cell's 27 stencil buckets in walk order
// (dz, dy, dx, then slot order inside a bucket), 64 per chunk, one candidate
// per lane, and tests each chunk against the run's particles in turn: for one
// particle the passing lanes are its next list entries in lane order, placed
// by a ballot and a prefix count. Every lane does useful work whatever the
// bucket sizes, and no lane loops over its own appends.
//
// A chunk's lanes find their bucket through marks: each non-empty bucket whose
// first candidate falls in the chunk writes its key (stencil index + 1 in the
// top bits, its entry offset below) at that lane in LDS, and an inclusive max
// over the lanes, carried from chunk to chunk, gives each lane the last bucket
// that opened at or before it.
// ---------------------------------------------------------------------------

// Inclusive max / sum over the wave's 64 lanes: row shifts inside each row of
// 16 lanes, then the row broadcasts of lanes 15 and 31. Lanes a step does not
// reach read 0, the identity of both.
template <int CTRL, int ROWS>
static __device__ __forceinline__ uint32_t dpp_or_zero(uint32_t v) {
  return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, CTRL, ROWS, 0xf, false);
}
static __device__ __forceinline__ uint32_t wave_inclusive_max64(uint32_t v) {
  v = max(v, dpp_or_zero<0x111, 0xf>(v));  // row_shr:1
  v = max(v, dpp_or_zero<0x112, 0xf>(v));  // row_shr:2
  v = max(v, dpp_or_zero<0x114, 0xf>(v));  // row_shr:4
  v = max(v, dpp_or_zero<0x118, 0xf>(v));  // row_shr:8
  v = max(v, dpp_or_zero<0x142, 0xa>(v));  // row_bcast:15 into rows 1 and 3
  v = max(v, dpp_or_zero<0x143, 0xc>(v));  // row_bcast:31 into rows 2 and 3
  return v;
}
static __device__ __forceinline__ uint32_t wave_inclusive_sum64(uint32_t v) {
  v += dpp_or_zero<0x111, 0xf>(v);
  v += dpp_or_zero<0x112, 0xf>(v);
  v += dpp_or_zero<0x114, 0xf>(v);
  v += dpp_or_zero<0x118, 0xf>(v);
  v += dpp_or_zero<0x142, 0xa>(v);
  v += dpp_or_zero<0x143, 0xc>(v);
  return v;
}

#define PF_RUN_CHUNKS 4  // chunks of 64 candidates held in registers at once

// Builds the lists of the wave's active particles. Every lane of the wave
// calls it. STAGED: candidates are the tile union's staged positions and the
// entries union indices, written as 16-bit halves of F.nl16's words; else
// candidates are read from xw and the entries are sorted slots in F.nl.
// WIDE: row offsets need full 32-bit products (n >= 2^22; n <= 2^24 keeps
// every offset within 32 bits). K: chunks of 64 candidates loaded at once.
// marks: 64 * K words of LDS owned
// by the wave. Returns the lane's list length; *over is set where a particle
// has more than PF_MAX_NEIGHBORS neighbours (its list keeps the first ones).
template <bool STAGED, bool WIDE, uint32_t K = PF_RUN_CHUNKS>
static __device__ __forceinline__ uint32_t walk_runs(const SceneDev& S, const FluidDev& F, const float4* __restrict__ xw,
                                     const float* u_px, const float* u_py, const float* u_pz,
                                     const uint32_t* h_key, const uint32_t* h_off, uint32_t* marks,
                                     bool active, bool leader, const int q[3], float3 xi, uint32_t s,
                                     bool* over) {
  const uint32_t lane = __lane_id();
  const uint32_t n = S.n;
  const float h2 = F.h2;
  const uint32_t s0 = __builtin_amdgcn_readfirstlane(s - lane);  // slot of lane 0
  const uint64_t actives = __builtin_amdgcn_ballot_w64(active);
  const uint64_t leaders = __builtin_amdgcn_ballot_w64(active && (leader || lane == 0u));
  // This lane's stencil offset (lanes 0..26, dz slowest, dx fastest).
  const uint32_t sdx = lane % 3u - 1u, sdy = (lane / 3u) % 3u - 1u, sdz = lane / 9u - 1u;
  PF_G uint32_t* __restrict__ nl = pf_g(F.nl);
  PF_G uint32_t* __restrict__ nl16 = pf_g(F.nl16);
  const PF_G float4* __restrict__ gxw = pf_g(xw);
  uint32_t count = 0;
  uint64_t overm = 0;
  uint64_t todo = leaders;
  while (todo != 0ull) {
    const uint32_t a = (uint32_t)__builtin_ctzll(todo);
    todo &= todo - 1ull;
    const uint64_t above = a == 63u ? 0ull : ~0ull << (a + 1u);
    const uint64_t stop = (leaders | ~actives) & above;
    const uint32_t end = stop ? (uint32_t)__builtin_ctzll(stop) : 64u;
    const uint64_t run = (end == 64u ? ~0ull : (1ull << end) - 1ull) & (~0ull << a);
    const uint32_t q0 = (uint32_t)__builtin_amdgcn_readlane(q[0], a);
    const uint32_t q1 = (uint32_t)__builtin_amdgcn_readlane(q[1], a);
    const uint32_t q2 = (uint32_t)__builtin_amdgcn_readlane(q[2], a);
    // The stencil buckets' entry ranges, their prefix over the walk order and
    // the marks' keys.
    uint32_t cn = 0, base = 0;
    if (lane < 27u) {
      const uint32_t b = pf_morton((int)(q0 + sdx), (int)(q1 + sdy), (int)(q2 + sdz)) & S.hash_mask;
      if (STAGED) {
        uint32_t h = tu_hash(b);
        for (uint32_t probe = 0; probe < PF_TU_HASH && h_key[h] != b + 1u; ++probe) h = (h + 1u) & (PF_TU_HASH - 1u);
        if (h_key[h] == b + 1u) {  // always: the cell's leader inserted b
          const uint32_t ho = h_off[h];
          base = ho >> 16;
          cn = ho & 0xffffu;
        }
      } else {
        base = S.start[b];
        cn = S.start[b + 1] - base;
      }
    }
    const uint32_t inc = wave_inclusive_sum64(cn);
    const uint32_t pre = inc - cn;
    const uint32_t C = (uint32_t)__builtin_amdgcn_readlane(inc, 63);
    const uint32_t key = cn ? ((lane + 1u) << 27) | (base - pre + (1u << 26)) : 0u;
    uint64_t live = run;  // particles of the run whose list is still open
    uint32_t carry = 0;
    for (uint32_t k0 = 0; k0 < C && live != 0ull; k0 += 64u * K) {
      float3 xj[K];
      uint32_t ej[K];
#pragma unroll
      for (uint32_t j = 0; j < K; ++j) {
        const uint32_t c0 = k0 + 64u * j;
        uint32_t* mk = marks + 64u * j;
        mk[lane] = 0u;
        __builtin_amdgcn_wave_barrier();
        if (key != 0u && pre - c0 < 64u) mk[pre - c0] = key;
        __builtin_amdgcn_wave_barrier();
        const uint32_t v = max(wave_inclusive_max64(mk[lane]), carry);
        carry = (uint32_t)__builtin_amdgcn_readlane(v, 63);
        const uint32_t kk = c0 + lane;
        const bool valid = kk < C;
        const uint32_t e = valid ? kk + (v & 0x7ffffffu) - (1u << 26) : 0u;
        ej[j] = e;
        float3 p;
        if (STAGED) {
          p = make_float3(u_px[e], u_py[e], u_pz[e]);
        } else {
          const float4 g = pf_f4(gxw[e]);
          p = make_float3(g.x, g.y, g.z);
        }
        // Lanes past the last candidate never pass the distance test.
        xj[j] = valid ? p : make_float3(__builtin_inff(), __builtin_inff(), __builtin_inff());
      }
      uint64_t rem = live;
      while (rem != 0ull) {
        const uint32_t i = (uint32_t)__builtin_ctzll(rem);
        rem &= rem - 1ull;
        const float sx = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(xi.x), i));
        const float sy = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(xi.y), i));
        const float sz = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(xi.z), i));
        uint32_t cnt = (uint32_t)__builtin_amdgcn_readlane(count, i);
        // The particle's list column, word rows n apart.
        PF_G char* const col = STAGED ? (PF_G char*)(nl16 + s0 + i) : (PF_G char*)(nl + s0 + i);
        // cnt counts every neighbour found; entries past PF_MAX_NEIGHBORS are
        // not stored, and a list that passes it closes after this block.
#pragma unroll
        for (uint32_t j = 0; j < K; ++j) {
          if (k0 + 64u * j >= C) break;
          const float3 d = make_float3(sx - xj[j].x, sy - xj[j].y, sz - xj[j].z);
          const bool hit = pf_dot3(d, d) < h2;
          const uint64_t m = __builtin_amdgcn_ballot_w64(hit);
          if (m != 0ull) {
            // The lane's entry: the particle's cnt-th plus the hits below it.
            const uint32_t pos = __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                                           __builtin_amdgcn_mbcnt_lo((uint32_t)m, cnt));
            if (hit & (pos < PF_MAX_NEIGHBORS)) {
              if (STAGED) {  // word row pos >> 1, half pos & 1
                const uint32_t off = (WIDE ? (pos >> 1) * (n * 4u) : __umul24(pos >> 1, n * 4u)) + ((pos & 1u) << 1);
                *(PF_G uint16_t*)(col + off) = (uint16_t)ej[j];
              } else {
                const uint32_t off = WIDE ? pos * (n * 4u) : __umul24(pos, n * 4u);
                *(PF_G uint32_t*)(col + off) = ej[j];
              }
            }
          }
          cnt += (uint32_t)__popcll(m);
        }
        count = lane == i ? min(cnt, (uint32_t)PF_MAX_NEIGHBORS) : count;
        if (cnt > PF_MAX_NEIGHBORS) {
          // The list is full and another neighbour exists: it overflows,
          // nothing later in the walk can enter it.
          live &= ~(1ull << i);
          overm |= 1ull << i;
        }
      }
    }
  }
  *over = ((overm >> lane) & 1ull) != 0ull;
  return active ? count : 0u;
}

// Surface tension's per-particle record, two float4s at F.normal + 2 s: the
// prediction and density, then the colour-field normal. A pair gathers its
// partner's record from one cache line.
static __device__ __forceinline__ void store_record(const SceneDev& S, const FluidDev& F, int p, uint32_t s,
                                                    float3 xi, float4 nrm) {
  F.normal[2 * s] = make_float4(xi.x, xi.y, xi.z, S.rho[p][s]);
  F.normal[2 * s + 1] = make_float4(nrm.x, nrm.y, nrm.z, 0.0f);
}

// Steps 1 and 2 of a tile's neighbour build. 1, the bucket set: the first
// particle of each run of equal query cells inserts its 27 stencil buckets in
// the LDS hash set h_key. 2, the union layout: buckets in hash-slot order,
// offsets by a block scan, into h_off ((union offset << 16) | count) and
// h_start (first sorted slot). *leader: the thread's particle opens a run.
// Returns the union size U, or PF_TU_GLOBAL when the set overflows (more
// than PF_TU_HASH_LIMIT buckets); uniform over the block. u_words holds the
// threads' query cells for the leader test (4 * PF_BLOCK words).
static __device__ __forceinline__ uint32_t nb_layout(const SceneDev& S, bool active, const int q[3], bool* leader,
                                                     uint32_t* h_key, uint32_t* h_off, uint32_t* h_start,
                                                     uint32_t* ctl, uint32_t* u_words) {
  const uint32_t tid = threadIdx.x;
  int (*q_sh)[PF_BLOCK] = reinterpret_cast<int (*)[PF_BLOCK]>(u_words);
  uint32_t* a_sh = u_words + 3 * PF_BLOCK;
  for (uint32_t i = tid; i < PF_TU_HASH; i += PF_BLOCK) h_key[i] = 0u;
  if (tid < 4) ctl[tid] = 0u;
  q_sh[0][tid] = q[0];
  q_sh[1][tid] = q[1];
  q_sh[2][tid] = q[2];
  a_sh[tid] = active ? 1 : 0;
  __syncthreads();
  *leader = active && (tid == 0 || !a_sh[tid - 1] || q_sh[0][tid - 1] != q[0] || q_sh[1][tid - 1] != q[1] ||
                       q_sh[2][tid - 1] != q[2]);
  // The leaders' cells, in lane order (h_start is free until the layout);
  // then every thread inserts (cell, stencil) pairs, 27 per leader.
  uint32_t nlead;
  const uint32_t li = block_exclusive_scan(*leader ? 1u : 0u, ctl + 4, &nlead);
  if (*leader) {
    h_start[3 * li + 0] = (uint32_t)q[0];
    h_start[3 * li + 1] = (uint32_t)q[1];
    h_start[3 * li + 2] = (uint32_t)q[2];
  }
  __syncthreads();
  {
    volatile uint32_t* give_up = &ctl[1];
    for (uint32_t task = tid; task < nlead * 27u && !*give_up; task += PF_BLOCK) {
      const uint32_t ci = task / 27u, d = task - ci * 27u;
      const uint32_t dx = d % 3u - 1u, dy = (d / 3u) % 3u - 1u, dz = d / 9u - 1u;
      const uint32_t b = pf_morton((int)(h_start[3 * ci + 0] + dx), (int)(h_start[3 * ci + 1] + dy),
                                   (int)(h_start[3 * ci + 2] + dz)) & S.hash_mask;
      uint32_t h = tu_hash(b);
      bool placed = false;
      for (uint32_t probe = 0; probe < PF_TU_HASH; ++probe) {
        const uint32_t prev = atomicCAS(&h_key[h], 0u, b + 1u);
        if (prev == 0u) {
          if (atomicAdd(&ctl[0], 1u) + 1u > PF_TU_HASH_LIMIT) *give_up = 1u;
          placed = true;
          break;
        }
        if (prev == b + 1u) {
          placed = true;
          break;
        }
        h = (h + 1u) & (PF_TU_HASH - 1u);
      }
      if (!placed) *give_up = 1u;
    }
  }
  __syncthreads();
  if (ctl[1] != 0u) return PF_TU_GLOBAL;
  constexpr int PER = PF_TU_HASH / PF_BLOCK;
  uint32_t cn[PER], st[PER], sum = 0;
#pragma unroll
  for (int i = 0; i < PER; ++i) {
    const uint32_t key = h_key[tid * PER + i];
    cn[i] = 0u;
    st[i] = 0u;
    if (key) {
      st[i] = S.start[key - 1u];
      cn[i] = S.start[key] - st[i];
    }
    sum += cn[i];
  }
  uint32_t U = 0;
  uint32_t off = block_exclusive_scan(sum, ctl + 4, &U);
#pragma unroll
  for (int i = 0; i < PER; ++i) {
    h_off[tid * PER + i] = (min(off, 0xffffu) << 16) | min(cn[i], 0xffffu);
    h_start[tid * PER + i] = st[i];
    off += cn[i];
  }
  __syncthreads();
  return U;
}

// Step 3, staging the union's U entries (at most PER * PF_BLOCK): entry u
// belongs to the last bucket whose union offset is not above u. Each bucket
// marks its offset with its hash slot + 1 and a chunked prefix max spreads the
// marks; the marks and the resolved slots use u_words (the positions' LDS,
// PER entries per thread). Positions go to u_px/u_py/u_pz, sorted slots to
// slots[u] (global or LDS). Every thread of the block calls it.
template <int PER>
static __device__ __forceinline__ void nb_stage(uint32_t U, const float4* __restrict__ xw, const uint32_t* h_off,
                                                const uint32_t* h_start, uint32_t* ctl, uint32_t* u_words,
                                                float* u_px, float* u_py, float* u_pz, uint32_t* slots) {
  static_assert(PER % 4 == 0, "nb_stage reads marks four at a time");
  const uint32_t tid = threadIdx.x;
  uint32_t* mark = u_words;
  const uint32_t uall = (U + PER - 1u) / PER * PER;
  for (uint32_t u = tid; u < uall; u += PF_BLOCK) mark[u] = 0u;
  __syncthreads();
  constexpr int HPER = PF_TU_HASH / PF_BLOCK;
#pragma unroll
  for (int i = 0; i < HPER; ++i) {
    const uint32_t h = tid * HPER + i;
    const uint32_t ho = h_off[h];
    if (ho & 0xffffu) mark[ho >> 16] = h + 1u;
  }
  __syncthreads();
  const uint32_t u0 = tid * (uint32_t)PER;
  uint32_t mk[PER], run = 0;
  if (u0 < U) {
#pragma unroll
    for (int k4 = 0; k4 < PER; k4 += 4) {
      const uint4 a = *reinterpret_cast<const uint4*>(mark + u0 + k4);
      const uint32_t m[4] = {a.x, a.y, a.z, a.w};
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        run = max(run, m[k]);
        mk[k4 + k] = run;
      }
    }
  }
  const uint32_t carry = block_exclusive_max(run, ctl + 4);
  if (u0 < U) {
#pragma unroll
    for (int k4 = 0; k4 < PER; k4 += 4) {
      uint32_t r[4];
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const uint32_t h = max(mk[k4 + k], carry) - 1u;  // entry 0 opens the first bucket
        r[k] = h_start[h] + ((u0 + (uint32_t)(k4 + k)) - (h_off[h] >> 16));
      }
      *reinterpret_cast<uint4*>(mark + u0 + k4) = make_uint4(r[0], r[1], r[2], r[3]);
    }
  }
  __syncthreads();
  uint32_t js[PER];
#pragma unroll
  for (int i = 0; i < PER; ++i) {
    const uint32_t u = tid + i * PF_BLOCK;
    js[i] = u < U ? mark[u] : 0u;
  }
  __syncthreads();
  float4 xv[PER];
#pragma unroll
  for (int i = 0; i < PER; ++i) xv[i] = xw[js[i]];
#pragma unroll
  for (int i = 0; i < PER; ++i) {
    const uint32_t u = tid + i * PF_BLOCK;
    if (u < U) {
      u_px[u] = xv[i].x;
      u_py[u] = xv[i].y;
      u_pz[u] = xv[i].z;
      slots[u] = js[i];
    }
  }
}

// ---------------------------------------------------------------------------
// Neighbour list (+ colour-field normal when surface tension is on)
//
// Each 256-slot tile first gathers the set of buckets its fluid particles
// walk (all 27 stencil buckets of each distinct query cell) in an LDS hash
// set, stages those buckets' particles in LDS in one pass (the tile union),
// and walks them there, one cell run per wave at a time (walk_runs). The list
// keeps the reference's order and cap (walk order dz->dy->dx, slot order
// inside a bucket, first 64 within h) and stores 16-bit union indices, two per
// word. The colour-field normal is summed over the finished list afterwards,
// in list order. A tile whose union does not fit (more than PF_TU_MAX
// particles or PF_TU_HASH_LIMIT buckets) walks global memory the same way and
// stores global slots, then compacts them into a union when the listed
// particles fit; every tile of a table smaller than 64 buckets (where stencil
// buckets can coincide) walks global memory one particle per lane with the
// reference's bucket dedup. All walks give the same list.
// With `snap` set, a contact stabilisation has moved some predictions since
// the sort: the query point is the current prediction (xs[c]) and candidates
// are tested at the positions the grid was built from (xs[c ^ 1], the
// reference's sortedPredictedPositions, neighborList.ts:59), while the
// colour-field normal reads current positions.
// ---------------------------------------------------------------------------
template <bool NORMAL, bool PERSIST>
__global__ __launch_bounds__(PF_BLOCK) void k_neighbors(const SceneDev* __restrict__ scenes,
                                                        const Tile* __restrict__ tiles,
                                                        uint32_t ntiles, int fi, int p, int c,
                                                        uint32_t cap, int snap) {
  __shared__ uint32_t h_tab[3 * PF_TU_HASH];
  uint32_t* h_key = h_tab;                   // bucket + 1, 0 = empty
  uint32_t* h_off = h_tab + PF_TU_HASH;      // (union offset << 16) | bucket count
  uint32_t* h_start = h_tab + 2 * PF_TU_HASH;  // first sorted slot of the bucket
  // The staged union's positions, x, y and z arrays of cap entries (dynamic
  // LDS, 12 bytes per entry and at least 4 KB); before staging the same words
  // hold each thread's query cell and activity for the leader test.
  extern __shared__ float4 pf_lds[];
  uint32_t* u_words = reinterpret_cast<uint32_t*>(pf_lds);
  float* u_px = reinterpret_cast<float*>(pf_lds);
  float* u_py = u_px + cap;
  float* u_pz = u_py + cap;
  __shared__ uint32_t ctl[4 + PF_BLOCK / 64];  // distinct buckets, give-up flag, -, -, scan
  pf_for_tiles<PERSIST>(ntiles, [&](uint32_t ti) {
    const Tile t = tiles[ti];
    const SceneDev& S = scenes[t.scene];
    const FluidDev& F = S.fluid[fi];
    const uint32_t tid = threadIdx.x;
    const uint32_t s = t.first + tid;
    const uint32_t tile = t.first / PF_BLOCK;
    const uint32_t n = S.n;
    const bool active = s < n && in_range(S.id[p][s], F.start, F.count);
    const float4* __restrict__ xs = S.xs[c];
    const float4* __restrict__ xw = snap ? S.xs[c ^ 1] : xs;  // candidate positions
    float3 xi = make_float3(0.f, 0.f, 0.f);
    int q[3] = {0, 0, 0};
    if (active) {
      xi = pf_xyz(xs[s]);
      const float r[3] = {(xi.x - S.ox) / S.cell, (xi.y - S.oy) / S.cell, (xi.z - S.oz) / S.cell};
  #pragma unroll
      for (int a = 0; a < 3; ++a) q[a] = pf_f2i_sat(floorf(r[a]));
    }
    const bool want_normal = NORMAL && (F.effects & PF_FX_SURFACE_TENSION);
    uint32_t count = 0;
    bool over = false;

    // Tables below 64 buckets: global walk with the reference's bucket dedup.
    // Such tables belong to scenes of at most 16 particles (H = nextPow2(2N)).
    if (S.H < 64u) {
      if (tid == 0) F.tu_size[tile] = PF_TU_GLOBAL;
      if (!active) return;
      uint32_t visited[27];
      uint32_t nvisited = 0;
      for (int dz = -1; dz <= 1 && !over; ++dz)
        for (int dy = -1; dy <= 1 && !over; ++dy)
          for (int dx = -1; dx <= 1 && !over; ++dx) {
            const uint32_t b = pf_morton((int)((uint32_t)q[0] + (uint32_t)dx), (int)((uint32_t)q[1] + (uint32_t)dy),
                                         (int)((uint32_t)q[2] + (uint32_t)dz)) & S.hash_mask;
            bool seen = false;
            for (uint32_t k = 0; k < nvisited; ++k) seen |= (visited[k] == b);
            visited[nvisited++] = b;
            if (seen) continue;
            for (uint32_t j = S.start[b]; j < S.start[b + 1] && !over; ++j) {
              const float4 xj = xw[j];
              const float3 d = make_float3(xi.x - xj.x, xi.y - xj.y, xi.z - xj.z);
              if (pf_dot3(d, d) < F.h2) {
                if (count < PF_MAX_NEIGHBORS) {
                  F.nl[(size_t)count * n + s] = j;
                  ++count;
                } else {
                  over = true;
                }
              }
            }
          }
      F.nl_count[s] = count;
      if (want_normal) store_record(S, F, p, s, xi, color_normal_global(S, F, p, xs, xi, s, count));
      if (over) flag_set(S, p ^ 1, PF_FLAG_NEIGHBORS);
      return;
    }

    bool leader = false;
    const uint32_t U = nb_layout(S, active, q, &leader, h_key, h_off, h_start, ctl, u_words);
    // A bucket set that overflowed (U == PF_TU_GLOBAL) or a union above the
    // capacity walks global memory.
    const bool global_walk = U > min(F.tu_limit, cap);

    // The run-walk marks, per wave: in h_start (free once the layout is read)
    // for the staged walk; in the union's LDS for the global walk, which keeps
    // the layout for its compaction.
    uint32_t* marks = h_start + (tid >> 6) * (64u * PF_RUN_CHUNKS);
    static_assert(PF_BLOCK / 64 * 64 * PF_RUN_CHUNKS <= PF_TU_HASH, "run-walk marks exceed h_start");
    if (global_walk) {
      // Global candidates arrive with a global load's latency: twice the
      // chunks in flight where the union's LDS holds their marks (from a
      // capacity of 683; 4 KB, the least it holds, fits the default chunks).
      const bool deep = cap * 3u >= (uint32_t)(PF_BLOCK / 64) * 64u * (2u * PF_RUN_CHUNKS);  // uniform
      uint32_t* gmarks = u_words + (tid >> 6) * (64u * (deep ? 2u * PF_RUN_CHUNKS : PF_RUN_CHUNKS));
      if (deep)
        count = n < (1u << 22) ? walk_runs<false, false, 2 * PF_RUN_CHUNKS>(S, F, xw, nullptr, nullptr, nullptr, h_key,
                                                                            h_off, gmarks, active, leader, q, xi, s, &over)
                               : walk_runs<false, true, 2 * PF_RUN_CHUNKS>(S, F, xw, nullptr, nullptr, nullptr, h_key,
                                                                           h_off, gmarks, active, leader, q, xi, s, &over);
      else
        count = n < (1u << 22) ? walk_runs<false, false>(S, F, xw, nullptr, nullptr, nullptr, h_key, h_off, gmarks, active,
                                                         leader, q, xi, s, &over)
                               : walk_runs<false, true>(S, F, xw, nullptr, nullptr, nullptr, h_key, h_off, gmarks, active,
                                                        leader, q, xi, s, &over);
      if (active) {
        F.nl_count[s] = count;
        if (want_normal) {
          // The lane's list was written by the wave's other lanes.
          __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
          __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
          store_record(S, F, p, s, xi, color_normal_global(S, F, p, xs, xi, s, count));
        }
        if (over) flag_set(S, p ^ 1, PF_FLAG_NEIGHBORS);
      }
      // The passes after this one only need the particles the lists name: if
      // those fit, they become the tile's union and the lists are rewritten as
      // 16-bit indices into it, so the tile is staged in LDS like the others.
      // A valid layout numbers them in layout order; a bucket set that
      // overflowed, or a layout too large for the LDS, sorts a hash table's.
      const uint32_t lds_words = (uint32_t)max((size_t)cap * 3u, (size_t)PF_BLOCK * 4u);
      const bool by_layout = U <= 0xffffu && compact_layout_words(U) <= lds_words;  // uniform
      // h_off and h_start (2 * PF_TU_HASH >= PF_TU_MAX words) sort the hash table's keys.
      static_assert(2 * PF_TU_HASH >= PF_TU_MAX, "compaction sort space");
      const uint32_t Uc =
          by_layout ? compact_layout(S, F, tile, h_key, h_off, h_start, U, u_words, min(F.tu_limit, cap), ctl,
                                     active ? count : 0u, s)
                    : compact_union(S, F, tile, u_words, cap, min(F.tu_limit, cap), ctl, active ? count : 0u, s, h_off);
      if (tid == 0) F.tu_size[tile] = Uc;
      return;
    }

    // 3. Stage the union.
    nb_stage<PF_TU_MAX / PF_BLOCK>(U, xw, h_off, h_start, ctl, u_words, u_px, u_py, u_pz,
                                   F.tu_slots + (size_t)tile * PF_TU_MAX);
    __syncthreads();

    // 4. Walk the union.
    PF_G uint32_t* __restrict__ nl16 = pf_g(F.nl16);
    count = n < (1u << 22)
                ? walk_runs<true, false>(S, F, xw, u_px, u_py, u_pz, h_key, h_off, marks, active, leader, q, xi, s, &over)
                : walk_runs<true, true>(S, F, xw, u_px, u_py, u_pz, h_key, h_off, marks, active, leader, q, xi, s, &over);
    if (active) {
      // The upper half of an odd list's last word is zero.
      if (count & 1u) *((PF_G uint16_t*)(nl16 + (size_t)(count >> 1) * n + s) + 1) = 0;
      F.nl_count[s] = count;
      if (want_normal) {
        // Colour-field normal over the finished list, in list order. The lane's
        // list was written by the wave's other lanes.
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
        const PF_G uint32_t* __restrict__ uslots = pf_g(F.tu_slots) + (size_t)tile * PF_TU_MAX;
        const PF_G float4* __restrict__ gxs = pf_g(xs);
        float3 nsum = make_float3(0.f, 0.f, 0.f);
        // Whole groups of four need no per-entry test; the last partial group does.
        auto group = [&](uint32_t k0, bool tail) {
          const uint32_t w0 = nl16[(size_t)(k0 >> 1) * n + s];
          const uint32_t w1 = nl16[(size_t)min((k0 >> 1) + 1u, (uint32_t)(PF_MAX_NEIGHBORS / 2 - 1)) * n + s];
          // entries past the count may hold stale indices: keep them inside the union
          const uint32_t ul = U ? U - 1u : 0u;
          const uint32_t idx[4] = {(uint32_t)min(w0 & (PF_TU_MAX - 1u), ul), (uint32_t)min((w0 >> 16) & (PF_TU_MAX - 1u), ul),
                                   (uint32_t)min(w1 & (PF_TU_MAX - 1u), ul), (uint32_t)min((w1 >> 16) & (PF_TU_MAX - 1u), ul)};
          float3 xj[4];
          float vol[4];
  #pragma unroll
          for (int u = 0; u < 4; ++u) {
            xj[u] = snap ? pf_xyz(gxs[uslots[idx[u]]]) : make_float3(u_px[idx[u]], u_py[idx[u]], u_pz[idx[u]]);
            vol[u] = S.has_boundary ? neighbor_volume(S, F, p, uslots[idx[u]]) : F.pv;
          }
  #pragma unroll
          for (int u = 0; u < 4; ++u) {
            const float3 d = make_float3(xi.x - xj[u].x, xi.y - xj[u].y, xi.z - xj[u].z);
            nsum = sel3(!tail || k0 + u < count, pf_add3(nsum, pf_mul3(d, pf_spiky_scale(d, F.h, F.spiky) * vol[u])), nsum);
          }
        };
        const uint32_t full = count & ~3u;
        for (uint32_t k0 = 0; k0 < full; k0 += 4) group(k0, false);
        if (full < count) group(full, true);
        store_record(S, F, p, s, xi, make_float4(nsum.x * F.h, nsum.y * F.h, nsum.z * F.h, 0.0f));
      }
      if (over) flag_set(S, p ^ 1, PF_FLAG_NEIGHBORS);
    }
    if (tid == 0) F.tu_size[tile] = U;
  });
}

// ---------------------------------------------------------------------------
// Neighbour-list cursors. Both walk the list four entries at a time with the
// next four entries' list words already loaded; rows are clamped so those
// loads stay in bounds past the end of the list, and entries past the count
// resolve to a harmless index whose terms the callers discard.
// ---------------------------------------------------------------------------
struct Nl16Cursor {  // staged tile: 16-bit union indices, two per word
  const PF_G uint32_t* __restrict__ col;
  size_t n;
  uint32_t w0, w1;
  uint32_t lim;  // last union index: entries past the count keep LDS reads inside the union
  __device__ __forceinline__ Nl16Cursor(const uint32_t* nl16, uint32_t nn, uint32_t s, uint32_t U)
      : col(pf_g(nl16) + s), n(nn), lim(U ? U - 1u : 0u) {
    w0 = col[0];
    w1 = col[n];
  }
  // Union indices of entries k0..k0+3; prefetches those of k0+4..k0+7.
  // CLAMP: some of the four may lie past the count (stale words); entries
  // inside the count are union indices already.
  template <bool CLAMP = true>
  __device__ __forceinline__ void next(uint32_t k0, uint32_t idx[4]) {
    const uint32_t c0 = w0, c1 = w1;
    const uint32_t r = min((k0 >> 1) + 2u, (uint32_t)(PF_MAX_NEIGHBORS / 2 - 2));
    w0 = col[(size_t)r * n];
    w1 = col[(size_t)(r + 1) * n];
    idx[0] = c0 & (PF_TU_MAX - 1);
    idx[1] = (c0 >> 16) & (PF_TU_MAX - 1);
    idx[2] = c1 & (PF_TU_MAX - 1);
    idx[3] = (c1 >> 16) & (PF_TU_MAX - 1);
    if (CLAMP) {
#pragma unroll
      for (int u = 0; u < 4; ++u) idx[u] = min(idx[u], lim);
    }
  }
};

// Calls body(std::true_type) when no active lane has a square root argument
// in the scaled range, else body(std::false_type); scaled is the OR of
// ballot(pf_sqrt_scaled(q)) over the group's arguments, so the choice is
// uniform over the wave.
template <class Body>
static __device__ __forceinline__ void with_sqrt_form(uint64_t scaled, Body&& body) {
  if (scaled == 0ull) body(std::true_type{});
  else body(std::false_type{});
}

struct NlCursor {  // global-gather tile: sorted slots
  const PF_G uint32_t* __restrict__ col;
  size_t n;
  uint32_t self, cnt;
  uint32_t w[4];
  __device__ __forceinline__ NlCursor(const uint32_t* nl, uint32_t nn, uint32_t s, uint32_t count)
      : col(pf_g(nl) + s), n(nn), self(s), cnt(count) {
#pragma unroll
    for (int u = 0; u < 4; ++u) w[u] = col[(size_t)u * n];
  }
  // Slots of entries k0..k0+3 (the own slot past the count).
  __device__ __forceinline__ void next(uint32_t k0, uint32_t jj[4]) {
#pragma unroll
    for (int u = 0; u < 4; ++u) jj[u] = (k0 + u < cnt) ? w[u] : self;
    const uint32_t r = min(k0 + 4u, (uint32_t)(PF_MAX_NEIGHBORS - 4));
#pragma unroll
    for (int u = 0; u < 4; ++u) w[u] = col[(size_t)(r + u) * n];
  }
};

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
  const float ir = pf_rcp_rn(r);  // r >= 1e-20
  const float3 coh = make_float3((off.x * ir) * cs, (off.y * ir) * cs, (off.z * ir) * cs);
  const float gm = F.gamma * F.mass;
  const float3 curv = make_float3(-((na.x - nb.x) * gm), -((na.y - nb.y) * gm), -((na.z - nb.z) * gm));
  *dv = make_float3((((coh.x + curv.x) * corr) * F.inv_mass) * dt, (((coh.y + curv.y) * corr) * F.inv_mass) * dt,
                    (((coh.z + curv.z) * corr) * F.inv_mass) * dt);
  return true;
}

// Adhesion spline A(r) (adhesion.ts:16-27).
static __device__ __forceinline__ float adhesion_spline(const FluidDev& F, float r) {
  const float root = pf_sqrt_rn(pf_sqrt_rn(fmaxf(((r * r) * F.adh_m4_over_h + r * 6.0f) - F.adh_2h, 0.0f)));
  return (r * 2.0f > F.h && r <= F.h) ? root * F.adh_coef : 0.0f;
}

// The partner's half of a pair, added to its accumulator: the change of the
// lower end negated before it is rounded to ticks, as the reference scatters
// it (cohesion.ts:96-100). Integer sums do not depend on the order the halves
// arrive in.
static __device__ __forceinline__ void scatter_half(int4* acc, uint32_t j, float3 dv, float scale) {
  PF_G int* aj = pf_g(reinterpret_cast<int*>(acc + j));
  pf_atomic_add_g(aj + 0, pf_ticks(-dv.x, scale));
  pf_atomic_add_g(aj + 1, pf_ticks(-dv.y, scale));
  pf_atomic_add_g(aj + 2, pf_ticks(-dv.z, scale));
}

// Adhesion toward one boundary neighbour of volume vol (adhesion.ts:34-60),
// added to the particle's own ticks.
static __device__ __forceinline__ void adhesion_add(const SceneDev& S, const FluidDev& F, float3 xi, float3 xj,
                                                    float vol, int tk[3]) {
  const float3 off = pf_sub3(xi, xj);
  const float r = pf_len3(off);
  if (r < 1e-20f) return;
  const float fs = -(((F.adhesion * F.mass) * (F.rho0 * vol)) * adhesion_spline(F, r));
  const float ir = pf_rcp_rn(r);  // r >= 1e-20
  const float3 f = make_float3((off.x * ir) * fs, (off.y * ir) * fs, (off.z * ir) * fs);
  tk[0] = pf_wrap_add(tk[0], pf_ticks((f.x * F.inv_mass) * S.dt, F.imp_scale));
  tk[1] = pf_wrap_add(tk[1], pf_ticks((f.y * F.inv_mass) * S.dt, F.imp_scale));
  tk[2] = pf_wrap_add(tk[2], pf_ticks((f.z * F.inv_mass) * S.dt, F.imp_scale));
}

// Surface tension and adhesion of one particle in a table below 64 buckets
// (global slots, one neighbour at a time; such scenes hold at most 16
// particles).
static __device__ void impulses_tiny(const SceneDev& S, const FluidDev& F, int p, int c, uint32_t s, uint32_t id,
                                     float3 xi, float3 ni, float rhoi, uint32_t cnt, bool st_on, bool adh_on,
                                     int tk[3]) {
  const float4* __restrict__ xs = S.xs[c];
  const uint32_t n = S.n;
  const uint32_t* __restrict__ nl = F.nl;
  const uint32_t fend = F.start + F.count;
  for (uint32_t k = 0; k < cnt; ++k) {
    const uint32_t j = nl[(size_t)k * n + s];
    if (st_on) {
      const uint32_t idj = S.id[p][j];
      if (idj > id && idj < fend) {
        float3 dv;
        if (pair_dv(F, S.dt, xi, pf_xyz(xs[j]), ni, pf_xyz(F.normal[2 * j + 1]), rhoi, S.rho[p][j], &dv)) {
          tk[0] = pf_wrap_add(tk[0], pf_ticks(dv.x, F.imp_scale));
          tk[1] = pf_wrap_add(tk[1], pf_ticks(dv.y, F.imp_scale));
          tk[2] = pf_wrap_add(tk[2], pf_ticks(dv.z, F.imp_scale));
          scatter_half(S.ct.acc, j, dv, F.imp_scale);
        }
      }
    }
    if (adh_on) {
      const float vol = S.bvol[p][j];
      if (vol > 0.0f) adhesion_add(S, F, xi, pf_xyz(xs[j]), vol, tk);
    }
  }
}

// Surface tension and adhesion of one particle over its list (cohesion.ts:63-103,
// adhesion.ts:34-60), four neighbours at a time with every gather of the four
// issued before the arithmetic. The reference scatters each pair from the
// list of its lower id, the change added to both ends, so a pair exists iff
// the lower end's list holds it: this particle evaluates the pairs its list
// holds with a higher fluid id, adds the change to its own ticks and the
// negated change straight to the partner's accumulator. Pairs where it is the
// upper end arrive from the partner the same way. (A staged tile with surface
// tension takes impulse_tile_flat instead.)
static __device__ __forceinline__ void impulse_walk(const SceneDev& S, const FluidDev& F, int p,
                                                    const float4* __restrict__ xs, uint32_t s, uint32_t id,
                                                    float3 xi, float3 ni, float rhoi, uint32_t cnt, bool st_pairs,
                                                    bool adh_on, uint32_t U, const uint32_t* __restrict__ uslots,
                                                    int tk[3]) {
  const uint32_t n = S.n;
  const uint32_t fend = F.start + F.count;
  const bool staged = U != PF_TU_GLOBAL;
  Nl16Cursor c16(F.nl16, n, s, staged ? U : 1u);
  NlCursor cg(F.nl, n, s, cnt);
  const PF_G uint32_t* __restrict__ g_slot = pf_g(uslots);
  const PF_G uint32_t* __restrict__ g_id = pf_g(S.id[p]);
  const PF_G float4* __restrict__ g_x = pf_g(xs);
  const PF_G float4* __restrict__ g_nrm = pf_g(F.normal);
  const PF_G float* __restrict__ g_rho = pf_g(S.rho[p]);
  const PF_G float* __restrict__ g_vol = pf_g(S.bvol[p]);
  for (uint32_t k0 = 0; k0 < cnt; k0 += 4) {
    uint32_t jj[4];
    if (staged) {
      uint32_t idx[4];
      c16.next(k0, idx);
#pragma unroll
      for (int u = 0; u < 4; ++u) jj[u] = g_slot[idx[u]];
    } else {
      cg.next(k0, jj);
    }
    uint32_t idj[4];
    float vol[4];
    bool lower[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      idj[u] = g_id[jj[u]];
      vol[u] = adh_on ? g_vol[jj[u]] : 0.0f;
    }
    float4 xj[4], nj[4];
    float rj[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const bool in = k0 + u < cnt;
      lower[u] = st_pairs && in && idj[u] > id && idj[u] < fend;
      xj[u] = make_float4(0.f, 0.f, 0.f, 0.f);
      nj[u] = make_float4(0.f, 0.f, 0.f, 0.f);
      rj[u] = 0.0f;
      if (lower[u] || (in && vol[u] > 0.0f)) xj[u] = pf_f4(g_x[jj[u]]);
      if (lower[u]) {
        nj[u] = pf_f4(g_nrm[2 * jj[u] + 1]);
        rj[u] = g_rho[jj[u]];
      }
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const float3 xjv = pf_xyz(xj[u]);
      if (lower[u]) {
        float3 dv;
        if (pair_dv(F, S.dt, xi, xjv, ni, pf_xyz(nj[u]), rhoi, rj[u], &dv)) {
          tk[0] = pf_wrap_add(tk[0], pf_ticks(dv.x, F.imp_scale));
          tk[1] = pf_wrap_add(tk[1], pf_ticks(dv.y, F.imp_scale));
          tk[2] = pf_wrap_add(tk[2], pf_ticks(dv.z, F.imp_scale));
          scatter_half(S.ct.acc, jj[u], dv, F.imp_scale);
        }
      }
      if (adh_on && k0 + u < cnt && vol[u] > 0.0f) adhesion_add(S, F, xi, xjv, vol[u], tk);
    }
  }
}

// One particle's surface tension and adhesion. With surface tension its own
// sum is parked, as bits, in its slot of the output prediction buffer for
// k_impulse_apply, which adds the halves its lower partners scattered; a
// fluid with adhesion alone applies here.
template <bool ST, bool ADH>
static __device__ __forceinline__ void impulse_particle(const SceneDev& S, const FluidDev& F, int p, int c,
                                                        uint32_t s, uint32_t tile) {
  const float4* __restrict__ xs = S.xs[c];
  const float4 xi4 = xs[s];
  const uint32_t id = S.id[p][s];
  if (!in_range(id, F.start, F.count)) {
    S.xs[c ^ 1][s] = xi4;
    return;
  }
  const float3 xi = pf_xyz(xi4);
  const uint32_t cnt = F.nl_count[s];
  int tk[3] = {0, 0, 0};
  float3 ni = make_float3(0.f, 0.f, 0.f);
  float rhoi = 0.0f;
  // A scene of a batch may lack an effect another scene has.
  const bool st_on = ST && (F.effects & PF_FX_SURFACE_TENSION);
  // Adhesion pulls only toward boundary particles (boundaryVolume > 0).
  const bool adh_on = ADH && (F.effects & PF_FX_ADHESION) && S.has_boundary;
  if (!(st_on || adh_on)) {
    S.xs[c ^ 1][s] = xi4;
    return;
  }
  if (st_on) {
    ni = pf_xyz(F.normal[2 * s + 1]);
    rhoi = S.rho[p][s];
  }
  const uint32_t U = F.tu_size[tile];
  const uint32_t* __restrict__ uslots = F.tu_slots + (size_t)tile * PF_TU_MAX;
  if (S.H < 64u) {
    impulses_tiny(S, F, p, c, s, id, xi, ni, rhoi, cnt, st_on, adh_on, tk);
  } else {
    impulse_walk(S, F, p, xs, s, id, xi, ni, rhoi, cnt, st_on, adh_on, U, uslots, tk);
  }
  if (st_on) {
    S.xs[c ^ 1][s] = make_float4(__int_as_float(tk[0]), __int_as_float(tk[1]), __int_as_float(tk[2]), xi4.w);
    return;
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

// Surface tension of a staged tile with its pairs spread over the wave. The
// lanes read their lists eight entries at a time (the next eight already in
// flight); each entry naming a pair the particle evaluates (a higher fluid id)
// goes to the wave's pair queue in LDS as (lane, union index), and the wave
// then evaluates the queue 64 pairs at a time, one per lane, whoever's pairs
// they are. The queue holds every lane's first pair, then every lane's second
// and so on, so the 64 pairs of a pass mostly belong to different particles
// and their own-sum atomics do not collide. Both halves of a pair are
// fixed-point sums (the particle's own in own_l, the partner's in acc_l), so
// the order the pairs are evaluated in does not change them. s_id and s_slot
// hold the union's particle ids and sorted slots: the lower-end test and a
// pair's partner slot read LDS only, and a pass's records (its own and its
// partner's, from global memory) are loaded a pass ahead. Adhesion stays with
// its particle's lane. Every thread of the block calls it.
#define PF_PAIR_QUEUE (64 * 8)  // a round's pairs per wave: eight entries per lane
template <bool ADH>
static __device__ __forceinline__ void impulse_tile_flat(const SceneDev& S, const FluidDev& F, int p, int c,
                                                         uint32_t first, uint32_t U, int* acc_l,
                                                         const uint32_t* s_id, const uint32_t* s_slot,
                                                         uint32_t* q_l, int* own_l) {
  const uint32_t lane = __lane_id();
  const uint32_t wave = threadIdx.x >> 6;
  const uint32_t s = first + threadIdx.x;
  const uint32_t n = S.n;
  const float4* __restrict__ xs = S.xs[c];
  const uint32_t fend = F.start + F.count;
  const uint32_t lim = U ? U - 1u : 0u;
  const bool adh_on = ADH && (F.effects & PF_FX_ADHESION) && S.has_boundary;
  uint32_t* q = q_l + wave * PF_PAIR_QUEUE;
  int* own = own_l + wave * 64 * 3;
  const uint32_t s0 = first + wave * 64u;  // slot of the wave's lane 0 (< n whenever a lane has a list)
  float4 xi4 = make_float4(0.f, 0.f, 0.f, 0.f);
  uint32_t id = 0, cnt = 0;
  bool act = false;
  if (s < n) {
    xi4 = xs[s];
    id = S.id[p][s];
    act = in_range(id, F.start, F.count);
  }
  if (act) cnt = F.nl_count[s];
  own[3 * lane + 0] = 0;
  own[3 * lane + 1] = 0;
  own[3 * lane + 2] = 0;
  int tk[3] = {0, 0, 0};  // adhesion
  const PF_G uint32_t* __restrict__ col = pf_g(F.nl16) + s;
  const PF_G float4* __restrict__ g_rec = pf_g(F.normal);
  const PF_G float4* __restrict__ g_x = pf_g(xs);
  const PF_G float* __restrict__ g_vol = pf_g(S.bvol[p]);
  const float3 xi = pf_xyz(xi4);
  constexpr uint32_t rmax = PF_MAX_NEIGHBORS / 2 - 1;
  // The list words of a round (entries k0..k0+7), rows clamped inside the list.
  auto load_words = [&](uint32_t k0, uint32_t w[4]) {
    const uint32_t r0 = k0 >> 1;
#pragma unroll
    for (uint32_t i = 0; i < 4; ++i) w[i] = k0 < cnt ? col[(size_t)min(r0 + i, rmax) * n] : 0u;
  };
  uint32_t wn[4];
  load_words(0, wn);
  for (uint32_t k0 = 0; __builtin_amdgcn_ballot_w64(k0 < cnt) != 0ull; k0 += 8u) {
    const uint32_t w0 = wn[0], w1 = wn[1], w2 = wn[2], w3 = wn[3];
    load_words(k0 + 8u, wn);  // the next round's words, in flight during this round
    uint32_t m = 0, ma = 0;
    auto entry = [=](uint32_t u) {  // union index of entry k0 + u
      const bool b1 = (u & 2u) != 0u, b2 = (u & 4u) != 0u;
      const uint32_t lo = b1 ? w1 : w0, hi = b1 ? w3 : w2;
      const uint32_t a = b2 ? hi : lo;
      return min((a >> ((u & 1u) * 16u)) & (PF_TU_MAX - 1u), lim);
    };
    if (k0 < cnt) {
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const uint32_t ix = entry((uint32_t)u);
        const bool in = k0 + (uint32_t)u < cnt;
        const uint32_t idj = s_id[ix];
        m |= (in && idj > id && idj < fend) ? (1u << u) : 0u;
        if (adh_on) ma |= (in && g_vol[s_slot[ix]] > 0.0f) ? (1u << u) : 0u;
      }
    }
    // The round's pairs, queued round-robin over the lanes: every lane's
    // first pair in lane order, then every second one, and so on.
    const uint32_t np = (uint32_t)__popc(m);
    uint32_t total = 0;
    {
      uint32_t mm = m;
#pragma unroll
      for (uint32_t r = 0; r < 8u; ++r) {
        const bool has = np > r;
        const uint64_t b = __builtin_amdgcn_ballot_w64(has);
        if (b == 0ull) break;
        if (has) {
          const uint32_t at = total + __builtin_amdgcn_mbcnt_hi((uint32_t)(b >> 32),
                                                                __builtin_amdgcn_mbcnt_lo((uint32_t)b, 0u));
          q[at] = (lane << 11) | entry((uint32_t)__builtin_ctz(mm));
          mm &= mm - 1u;
        }
        total += (uint32_t)__popcll(b);
      }
    }
    static_assert(PF_TU_MAX <= (1u << 11), "union index must fit 11 bits");
    __builtin_amdgcn_wave_barrier();
    // 64 pairs per pass, the next pass's records loaded before this one's
    // arithmetic. Lanes past the queue's end read pair 0's entry (union index
    // 0 exists: a list holds its own particle).
    struct Pass {
      uint32_t ix, il;
      float4 xa, na, xb, nb;
    };
    auto load_pass = [&](uint32_t qi) {
      Pass d;
      const uint32_t e = qi < total ? q[qi] : 0u;
      d.ix = e & 2047u;
      d.il = e >> 11;
      const uint32_t j = s_slot[d.ix];
      const uint32_t so = min(s0 + d.il, n - 1u);
      d.xa = pf_f4(g_rec[2 * so]);
      d.na = pf_f4(g_rec[2 * so + 1]);
      d.xb = pf_f4(g_rec[2 * j]);
      d.nb = pf_f4(g_rec[2 * j + 1]);
      return d;
    };
    Pass cur = load_pass(lane);
    for (uint32_t q0 = 0; q0 < total; q0 += 64u) {
      const bool mine = q0 + lane < total;
      const Pass nxt = load_pass(q0 + 64u + lane);
      float3 dv;
      if (mine && pair_dv(F, S.dt, pf_xyz(cur.xa), pf_xyz(cur.xb), pf_xyz(cur.na), pf_xyz(cur.nb), cur.xa.w,
                          cur.xb.w, &dv)) {
        atomicAdd(own + 3 * cur.il + 0, pf_ticks(dv.x, F.imp_scale));
        atomicAdd(own + 3 * cur.il + 1, pf_ticks(dv.y, F.imp_scale));
        atomicAdd(own + 3 * cur.il + 2, pf_ticks(dv.z, F.imp_scale));
        atomicAdd(acc_l + 3 * cur.ix + 0, pf_ticks(-dv.x, F.imp_scale));
        atomicAdd(acc_l + 3 * cur.ix + 1, pf_ticks(-dv.y, F.imp_scale));
        atomicAdd(acc_l + 3 * cur.ix + 2, pf_ticks(-dv.z, F.imp_scale));
      }
      cur = nxt;
    }
    while (ma != 0u) {
      const uint32_t jj = s_slot[entry((uint32_t)__builtin_ctz(ma))];
      ma &= ma - 1u;
      adhesion_add(S, F, xi, pf_xyz(pf_f4(g_x[jj])), g_vol[jj], tk);
    }
    // The next round's queue overwrites this one's.
    __builtin_amdgcn_wave_barrier();
  }
  __builtin_amdgcn_wave_barrier();
  if (s >= n) return;
  if (!act) {
    S.xs[c ^ 1][s] = xi4;
    return;
  }
  // The particle's own sum, parked as bits for k_impulse_apply.
  tk[0] = pf_wrap_add(tk[0], own[3 * lane + 0]);
  tk[1] = pf_wrap_add(tk[1], own[3 * lane + 1]);
  tk[2] = pf_wrap_add(tk[2], own[3 * lane + 2]);
  S.xs[c ^ 1][s] = make_float4(__int_as_float(tk[0]), __int_as_float(tk[1]), __int_as_float(tk[2]), xi4.w);
}

// A staged tile sums the halves its particles scatter to its union's entries
// in LDS and adds each touched entry's sums to the accumulator once at the
// end; k_impulse_apply applies once every half has arrived.
template <bool ST, bool ADH, bool PERSIST>
__global__ __launch_bounds__(PF_BLOCK) void k_impulses(const SceneDev* __restrict__ scenes,
                                                       const Tile* __restrict__ tiles,
                                                       uint32_t ntiles, int fi, int p, int c) {
  __shared__ int acc_l[ST ? 3 * PF_TU_MAX : 1];
  __shared__ uint32_t s_id[ST ? PF_TU_MAX : 1];
  __shared__ uint32_t s_slot[ST ? PF_TU_MAX : 1];
  // The waves' pair queues and own sums.
  __shared__ uint32_t q_l[ST ? PF_BLOCK / 64 * PF_PAIR_QUEUE : 1];
  __shared__ int own_l[ST ? PF_BLOCK * 3 : 1];
  pf_for_tiles<PERSIST>(ntiles, [&](uint32_t ti) {
    const Tile t = tiles[ti];
    const SceneDev& S = scenes[t.scene];
    const FluidDev& F = S.fluid[fi];
    const uint32_t tile = t.first / PF_BLOCK;
    // Uniform over the block.
    const bool st_on = ST && (F.effects & PF_FX_SURFACE_TENSION);
    const uint32_t U = st_on ? F.tu_size[tile] : PF_TU_GLOBAL;
    const bool lds = ST && st_on && S.H >= 64u && U != PF_TU_GLOBAL;
    if (lds) {
      for (uint32_t u = threadIdx.x; u < 3 * U; u += PF_BLOCK) acc_l[u] = 0;
      // The union's particle ids, every gather of a thread in flight at once.
      const PF_G uint32_t* __restrict__ g_id = pf_g(S.id[p]);
      uint32_t js[PF_TU_PER], ids[PF_TU_PER];
      tu_load_slots(F.tu_slots + (size_t)tile * PF_TU_MAX, U, js);
  #pragma unroll
      for (int i = 0; i < PF_TU_PER; ++i) ids[i] = g_id[js[i]];
  #pragma unroll
      for (int i = 0; i < PF_TU_PER; ++i) {
        const uint32_t u = threadIdx.x + i * PF_BLOCK;
        if (u < U) {
          s_id[u] = ids[i];
          s_slot[u] = js[i];
        }
      }
      __syncthreads();
    }
    const uint32_t s = t.first + threadIdx.x;
    if (lds) {
      if constexpr (ST) impulse_tile_flat<ADH>(S, F, p, c, t.first, U, acc_l, s_id, s_slot, q_l, own_l);
    } else if (s < S.n) {
      impulse_particle<ST, ADH>(S, F, p, c, s, tile);
    }
    if (lds) {
      __syncthreads();
      for (uint32_t u = threadIdx.x; u < U; u += PF_BLOCK) {
        const int a0 = acc_l[3 * u], a1 = acc_l[3 * u + 1], a2 = acc_l[3 * u + 2];
        if ((a0 | a1 | a2) != 0) {
          PF_G int* a = pf_g(reinterpret_cast<int*>(S.ct.acc + s_slot[u]));
          pf_atomic_add_g(a + 0, a0);
          pf_atomic_add_g(a + 1, a1);
          pf_atomic_add_g(a + 2, a2);
        }
      }
    }
  });
}

// The impulse apply (FluidSystem.ts:253-258, accumulator.ts:101-131) of the
// fluid particles of scenes with surface tension: the particle's own sum
// (parked in its output prediction slot) plus the halves scattered to it;
// v += change, x* += change dt, sums cleared.
template <bool PERSIST>
__global__ __launch_bounds__(PF_BLOCK) void k_impulse_apply(const SceneDev* __restrict__ scenes,
                                                            const Tile* __restrict__ tiles, uint32_t ntiles,
                                                            int fi, int p, int c) {
  pf_for_tiles<PERSIST>(ntiles, [&](uint32_t ti) {
    const Tile t = tiles[ti];
    const SceneDev& S = scenes[t.scene];
    const FluidDev& F = S.fluid[fi];
    if (!(F.effects & PF_FX_SURFACE_TENSION)) return;
    const uint32_t s = t.first + threadIdx.x;
    if (s >= S.n || !in_range(S.id[p][s], F.start, F.count)) return;
    int* a = reinterpret_cast<int*>(S.ct.acc + s);
    const float4 own = S.xs[c ^ 1][s];
    int tk[3] = {pf_wrap_add(a[0], __float_as_int(own.x)), pf_wrap_add(a[1], __float_as_int(own.y)),
                 pf_wrap_add(a[2], __float_as_int(own.z))};
    a[0] = 0;
    a[1] = 0;
    a[2] = 0;
    bool sat;
    const float3 ch = ticks_to_change(tk, F.imp_inv, &sat);
    if (sat) flag_set(S, 0, PF_FLAG_IMPULSES);
    const float4 xi4 = S.xs[c][s];
    float4 v = S.v[p][s];
    v.x = v.x + ch.x;
    v.y = v.y + ch.y;
    v.z = v.z + ch.z;
    S.v[p][s] = v;
    S.xs[c ^ 1][s] = make_float4(xi4.x + ch.x * S.dt, xi4.y + ch.y * S.dt, xi4.z + ch.z * S.dt, xi4.w);
  });
}

// ---------------------------------------------------------------------------
// ViscositySolver (ViscositySolver.ts:78-123): Jacobi sweeps of
// (I - dt nu L) v = v0 over the fluid's neighbour lists, after its velocity
// passes. The iterates live in slot order for one substep; slots outside the
// fluid hold zero velocity in both, as the reference's id-indexed buffers do.
// ---------------------------------------------------------------------------
template <bool PERSIST>
__global__ __launch_bounds__(PF_BLOCK) void k_isv_seed(const SceneDev* __restrict__ scenes,
                                                       const Tile* __restrict__ tiles, uint32_t ntiles,
                                                       int fi, int p) {
  pf_for_tiles<PERSIST>(ntiles, [&](uint32_t ti) {
    const Tile t = tiles[ti];
    const SceneDev& S = scenes[t.scene];
    const FluidDev& F = S.fluid[fi];
    const uint32_t s = t.first + threadIdx.x;
    if (s >= S.n || F.isv_iters == 0u) return;
    if (in_range(S.id[p][s], F.start, F.count)) {
      const float4 v = S.v[p][s];
      F.isv_rhs[s] = v;
      F.isv_a[s] = v;
    } else {
      F.isv_a[s] = make_float4(0.f, 0.f, 0.f, 0.f);
      F.isv_b[s] = make_float4(0.f, 0.f, 0.f, 0.f);
    }
  });
}

// Sweep k reads iterate k & 1 and writes the other. A scene of a batch with
// fewer sweeps copies its iterate through the extra ones.
template <bool PERSIST>
__global__ __launch_bounds__(PF_BLOCK) void k_isv_sweep(const SceneDev* __restrict__ scenes,
                                                        const Tile* __restrict__ tiles, uint32_t ntiles,
                                                        int fi, int p, uint32_t k) {
  pf_for_tiles<PERSIST>(ntiles, [&](uint32_t ti) {
    const Tile t = tiles[ti];
    const SceneDev& S = scenes[t.scene];
    const FluidDev& F = S.fluid[fi];
    const uint32_t s = t.first + threadIdx.x;
    const uint32_t n = S.n;
    if (s >= n || F.isv_iters == 0u || !in_range(S.id[p][s], F.start, F.count)) return;
    const float4* __restrict__ src = (k & 1u) ? F.isv_b : F.isv_a;
    float4* __restrict__ dst = (k & 1u) ? F.isv_a : F.isv_b;
    if (k >= F.isv_iters) {
      dst[s] = src[s];
      return;
    }
    const float4* __restrict__ X = S.x[p];
    const float3 xi = pf_xyz(X[s]);
    const uint32_t cnt = F.nl_count[s];
    const uint32_t tile = t.first / PF_BLOCK;
    const uint32_t U = F.tu_size[tile];
    const bool staged = U != PF_TU_GLOBAL;
    const uint32_t* __restrict__ uslots = F.tu_slots + (size_t)tile * PF_TU_MAX;
    Nl16Cursor c16(F.nl16, n, s, staged ? U : 1u);
    NlCursor cg(F.nl, n, s, cnt);
    float3 sum = make_float3(0.f, 0.f, 0.f);
    float weight = 0.0f;
    for (uint32_t k0 = 0; k0 < cnt; k0 += 4) {
      uint32_t jj[4];
      if (staged) {
        uint32_t idx[4];
        c16.next(k0, idx);
  #pragma unroll
        for (int u = 0; u < 4; ++u) jj[u] = uslots[idx[u]];
      } else {
        cg.next(k0, jj);
      }
      float4 xj[4], vj[4];
  #pragma unroll
      for (int u = 0; u < 4; ++u) {
        xj[u] = X[jj[u]];
        vj[u] = src[jj[u]];
      }
  #pragma unroll
      for (int u = 0; u < 4; ++u) {
        const float3 d = make_float3(xi.x - xj[u].x, xi.y - xj[u].y, xi.z - xj[u].z);
        const float w = pf_poly6(pf_dot3(d, d), F.h2, F.poly6) * F.pv;
        const bool ok = k0 + u < cnt;
        sum = sel3(ok, make_float3(sum.x + vj[u].x * w, sum.y + vj[u].y * w, sum.z + vj[u].z * w), sum);
        weight = ok ? weight + w : weight;
      }
    }
    const float rate = (F.isv_nu * S.dt) * F.isv_c10;
    const float4 r = F.isv_rhs[s];
    const float den = weight * rate + 1.0f;
    dst[s] = make_float4((r.x + sum.x * rate) / den, (r.y + sum.y * rate) / den, (r.z + sum.z * rate) / den, 0.0f);
  });
}

// The last iterate replaces the fluid's velocities (w kept).
template <bool PERSIST>
__global__ __launch_bounds__(PF_BLOCK) void k_isv_apply(const SceneDev* __restrict__ scenes,
                                                        const Tile* __restrict__ tiles, uint32_t ntiles,
                                                        int fi, int p, uint32_t sweeps) {
  pf_for_tiles<PERSIST>(ntiles, [&](uint32_t ti) {
    const Tile t = tiles[ti];
    const SceneDev& S = scenes[t.scene];
    const FluidDev& F = S.fluid[fi];
    const uint32_t s = t.first + threadIdx.x;
    if (s >= S.n || F.isv_iters == 0u || !in_range(S.id[p][s], F.start, F.count)) return;
    const float4 r = ((sweeps & 1u) ? F.isv_b : F.isv_a)[s];
    const float4 v = S.v[p][s];
    S.v[p][s] = make_float4(r.x, r.y, r.z, v.w);
  });
}

// ---------------------------------------------------------------------------
// Density and lambda; lambda is written to xs[c].w for the correction pass
// ---------------------------------------------------------------------------

// One neighbour's terms of the density and lambda sums (lambda.ts:40-58),
// added when ok.
template <bool DIRECT = false>
static __device__ __forceinline__ void lambda_term(const FluidDev& F, bool ok, float3 xi, float4 xj,
                                                   float mj, float wj, float& rho, float3& gs,
                                                   float& gss) {
  const float3 d = make_float3(xi.x - xj.x, xi.y - xj.y, xi.z - xj.z);
  const float q = pf_dot3(d, d);
  const float r1 = rho + mj * pf_poly6(q, F.h2, F.poly6);
  const float sc = pf_spiky_scale<DIRECT>(d, F.h, F.spiky) * mj;
  const float3 g = pf_mul3(d, sc);
  const float3 gs1 = pf_add3(gs, g);
  const float gss1 = gss + ((q * sc) * sc) * wj;  // |g|^2 = |d|^2 sc^2
  rho = ok ? r1 : rho;
  gs = sel3(ok, gs1, gs);
  gss = ok ? gss1 : gss;
}

template <bool PERSIST>
__global__ __launch_bounds__(PF_BLOCK) void k_lambda(const SceneDev* __restrict__ scenes,
                                                     const Tile* __restrict__ tiles, uint32_t ntiles,
                                                     int fi, int p, int c, uint32_t cap) {
  extern __shared__ float4 pf_lds[];
  float4* u_x = pf_lds;                                  // position, neighbour mass
  float* u_w = reinterpret_cast<float*>(pf_lds + cap);   // inverse mass
  pf_for_tiles<PERSIST>(ntiles, [&](uint32_t ti) {
    const Tile t = tiles[ti];
    const SceneDev& S = scenes[t.scene];
    const FluidDev& F = S.fluid[fi];
    const uint32_t s = t.first + threadIdx.x;
    const uint32_t tile = t.first / PF_BLOCK;
    const uint32_t n = S.n;
    float4* __restrict__ xs = S.xs[c];
    const uint32_t U = F.tu_size[tile];
    const bool staged = U != PF_TU_GLOBAL;
    if (staged) {
      uint32_t js[PF_TU_PER];
      tu_load_slots(F.tu_slots + (size_t)tile * PF_TU_MAX, U, js);
      float4 xv[PF_TU_PER];
      float mv[PF_TU_PER], wv[PF_TU_PER];
  #pragma unroll
      for (int i = 0; i < PF_TU_PER; ++i) {
        xv[i] = xs[js[i]];
        mv[i] = S.mass[js[i]];
        wv[i] = S.w[p][js[i]];
      }
      if (S.has_boundary) {
  #pragma unroll
        for (int i = 0; i < PF_TU_PER; ++i) mv[i] = neighbor_mass(S, F, p, js[i]);
      }
  #pragma unroll
      for (int i = 0; i < PF_TU_PER; ++i) {
        const uint32_t u = threadIdx.x + i * PF_BLOCK;
        if (u < U) {
          u_x[u] = make_float4(xv[i].x, xv[i].y, xv[i].z, mv[i]);
          u_w[u] = wv[i];
        }
      }
      __syncthreads();
    }
    if (s >= n || !in_range(S.id[p][s], F.start, F.count)) return;
    const float3 xi = pf_xyz(xs[s]);
    const uint32_t cnt = F.nl_count[s];
    float rho = 0.0f, gss = 0.0f;
    float3 gs = make_float3(0.f, 0.f, 0.f);
    // Four neighbours' loads in flight, then the sums in list order.
    if (staged) {
      Nl16Cursor cur(F.nl16, n, s, U);
      // Whole groups of four need no per-entry test; the last partial group does.
      auto group = [&](uint32_t k0, bool tail) {
        uint32_t idx[4];
        if (tail) cur.next<true>(k0, idx);
        else cur.next<false>(k0, idx);
        float4 xj[4];
        float wj[4];
        uint64_t scaled = 0;
  #pragma unroll
        for (int u = 0; u < 4; ++u) {
          xj[u] = u_x[idx[u]];
          wj[u] = u_w[idx[u]];
          const float3 d = make_float3(xi.x - xj[u].x, xi.y - xj[u].y, xi.z - xj[u].z);
          scaled |= __builtin_amdgcn_ballot_w64(pf_sqrt_scaled(pf_dot3(d, d)));
        }
        with_sqrt_form(scaled, [&](auto direct) {
  #pragma unroll
          for (int u = 0; u < 4; ++u)
            lambda_term<decltype(direct)::value>(F, !tail || k0 + u < cnt, xi, xj[u], xj[u].w, wj[u], rho, gs, gss);
        });
      };
      const uint32_t full = cnt & ~3u;
      for (uint32_t k0 = 0; k0 < full; k0 += 4) group(k0, false);
      if (full < cnt) group(full, true);
    } else {
      NlCursor cur(F.nl, n, s, cnt);
      for (uint32_t k0 = 0; k0 < cnt; k0 += 4) {
        uint32_t jj[4];
        cur.next(k0, jj);
        float4 xj[4];
        float mj[4], wj[4];
  #pragma unroll
        for (int u = 0; u < 4; ++u) {
          xj[u] = xs[jj[u]];
          mj[u] = neighbor_mass(S, F, p, jj[u]);
          wj[u] = S.w[p][jj[u]];
        }
  #pragma unroll
        for (int u = 0; u < 4; ++u) lambda_term(F, k0 + u < cnt, xi, xj[u], mj[u], wj[u], rho, gs, gss);
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
  });
}

// ---------------------------------------------------------------------------
// Position correction (+ container contact, accumulator apply, advect)
// ---------------------------------------------------------------------------

// One neighbour's term of the correction sum (positionDelta.ts:44-52), added
// when ok; returns the kernel gradient for the boundary reaction.
static __device__ __forceinline__ float3 delta_term(const FluidDev& F, bool ok, float3 xn, float li,
                                                    float4 xj, float mj, float3& sum) {
  const float3 d = make_float3(xn.x - xj.x, xn.y - xj.y, xn.z - xj.z);
  const float sp = pf_spiky_scale(d, F.h, F.spiky);
  sum = sel3(ok, pf_add3(sum, pf_mul3(d, sp * ((li + xj.w) * mj))), sum);
  return pf_mul3(d, sp);
}

// The matching push on a dynamic boundary neighbour j of this fluid
// (positionDelta.ts:54-69): -(lambda_i / rho0) * psi_j * w_j * grad W, into
// the reaction accumulator (Accumulator(10)) of slot j.
static __device__ __forceinline__ void reaction_add(const SceneDev& S, const FluidDev& F, int p, uint32_t j,
                                                    float li, float3 g) {
  const float bj = S.bvol[p][j];
  const float wj = S.w[p][j];
  if (!(bj > 0.0f && wj > 0.0f)) return;
  const uint32_t idj = S.id[p][j];
  bool inb = false;
  for (uint32_t b = 0; b < F.nboundary; ++b) inb = inb || in_range(idj, F.bstart[b], F.bcount[b]);
  if (!inb) return;
  const float rs = ((-(li / F.rho0)) * (F.rho0 * bj)) * wj;
  atomicAdd(F.reaction + 3 * (size_t)j + 0, pf_ticks(g.x * rs, F.rea_scale));
  atomicAdd(F.reaction + 3 * (size_t)j + 1, pf_ticks(g.y * rs, F.rea_scale));
  atomicAdd(F.reaction + 3 * (size_t)j + 2, pf_ticks(g.z * rs, F.rea_scale));
}

template <bool WALLS, bool ADVECT, bool PERSIST>
__global__ __launch_bounds__(PF_BLOCK) void k_delta(const SceneDev* __restrict__ scenes,
                                                    const Tile* __restrict__ tiles, uint32_t ntiles,
                                                    int fi, int p, int c, int first_iteration,
                                                    int substep, uint32_t cap) {
  extern __shared__ float4 pf_lds[];
  float4* u_x = pf_lds;                                  // position, lambda
  float* u_m = reinterpret_cast<float*>(pf_lds + cap);   // neighbour mass
  pf_for_tiles<PERSIST>(ntiles, [&](uint32_t ti) {
    const Tile t = tiles[ti];
    const SceneDev& S = scenes[t.scene];
    const FluidDev& F = S.fluid[fi];
    const uint32_t s = t.first + threadIdx.x;
    const uint32_t tile = t.first / PF_BLOCK;
    const uint32_t n = S.n;
    const float4* __restrict__ xs = S.xs[c];
    const uint32_t U = F.tu_size[tile];
    const bool staged = U != PF_TU_GLOBAL;
    if (staged) {
      uint32_t js[PF_TU_PER];
      tu_load_slots(F.tu_slots + (size_t)tile * PF_TU_MAX, U, js);
      float4 xv[PF_TU_PER];
      float mv[PF_TU_PER];
  #pragma unroll
      for (int i = 0; i < PF_TU_PER; ++i) {
        xv[i] = xs[js[i]];
        mv[i] = S.mass[js[i]];
      }
      if (S.has_boundary) {
  #pragma unroll
        for (int i = 0; i < PF_TU_PER; ++i) mv[i] = neighbor_mass(S, F, p, js[i]);
      }
  #pragma unroll
      for (int i = 0; i < PF_TU_PER; ++i) {
        const uint32_t u = threadIdx.x + i * PF_BLOCK;
        if (u < U) {
          u_x[u] = pf_f4(xv[i]);
          u_m[u] = mv[i];
        }
      }
      __syncthreads();
    }
    if (s >= n) return;
    const float4 xi4 = xs[s];
    float3 xn = pf_xyz(xi4);
    const float w = S.w[p][s];
    if (in_range(S.id[p][s], F.start, F.count)) {
      const uint32_t cnt = F.nl_count[s];
      const float li = xi4.w;
      const bool react = F.nboundary != 0u && li != 0.0f;
      const uint32_t* __restrict__ uslots = F.tu_slots + (size_t)tile * PF_TU_MAX;
      float3 sum = make_float3(0.f, 0.f, 0.f);
      if (staged) {
        Nl16Cursor cur(F.nl16, n, s, U);
        // Whole groups of four need no per-entry test; the last partial group
        // does. The boundary reaction has a walk of its own, so the common walk
        // carries no per-neighbour branch.
        auto group = [&](uint32_t k0, bool tail, bool with_reaction) {
          uint32_t idx[4];
          cur.next(k0, idx);
          float4 xj[4];
          float mj[4];
  #pragma unroll
          for (int u = 0; u < 4; ++u) {
            xj[u] = u_x[idx[u]];
            mj[u] = u_m[idx[u]];
          }
  #pragma unroll
          for (int u = 0; u < 4; ++u) {
            const bool ok = !tail || k0 + u < cnt;
            const float3 g = delta_term(F, ok, xn, li, xj[u], mj[u], sum);
            if (with_reaction && ok) reaction_add(S, F, p, uslots[idx[u]], li, g);
          }
        };
        const uint32_t full = cnt & ~3u;
        if (react) {
          for (uint32_t k0 = 0; k0 < full; k0 += 4) group(k0, false, true);
          if (full < cnt) group(full, true, true);
        } else {
          for (uint32_t k0 = 0; k0 < full; k0 += 4) group(k0, false, false);
          if (full < cnt) group(full, true, false);
        }
      } else {
        NlCursor cur(F.nl, n, s, cnt);
        for (uint32_t k0 = 0; k0 < cnt; k0 += 4) {
          uint32_t jj[4];
          cur.next(k0, jj);
          float4 xj[4];
          float mj[4];
  #pragma unroll
          for (int u = 0; u < 4; ++u) {
            xj[u] = xs[jj[u]];
            mj[u] = neighbor_mass(S, F, p, jj[u]);
          }
  #pragma unroll
          for (int u = 0; u < 4; ++u) {
            const float3 g = delta_term(F, k0 + u < cnt, xn, li, xj[u], mj[u], sum);
            if (react && k0 + u < cnt) reaction_add(S, F, p, jj[u], li, g);
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
        sdf_solve(S, s, xn, pf_xyz(x), w, first_iteration != 0, tk);
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
  });
}

// Apply and clear the boundary reaction over the span covering the fluid's
// boundary ranges (accumulator.ts:101-131 over coveringRange, FluidSystem.ts:280-287).
template <bool PERSIST>
__global__ __launch_bounds__(PF_BLOCK) void k_reaction_apply(const SceneDev* __restrict__ scenes,
                                                             const Tile* __restrict__ tiles,
                                                             uint32_t ntiles, int fi, int p, int c) {
  pf_for_tiles<PERSIST>(ntiles, [&](uint32_t ti) {
    const Tile t = tiles[ti];
    const SceneDev& S = scenes[t.scene];
    const FluidDev& F = S.fluid[fi];
    const uint32_t s = t.first + threadIdx.x;
    if (s >= S.n || F.nboundary == 0u) return;
    const uint32_t id = S.id[p][s];
    if (!(id >= F.rcover_lo && id < F.rcover_hi)) return;
    int* acc = F.reaction + 3 * (size_t)s;
    const int tk[3] = {acc[0], acc[1], acc[2]};
    bool sat;
    const float3 ch = ticks_to_change(tk, F.rea_inv, &sat);
    if (sat) flag_set(S, 0, PF_FLAG_REACTION);
    float4 x = S.xs[c][s];
    S.xs[c][s] = make_float4(x.x + ch.x, x.y + ch.y, x.z + ch.z, x.w);
    acc[0] = 0;
    acc[1] = 0;
    acc[2] = 0;
  });
}

// Akinci boundary volume V_i = 1 / max(sum_k W(x_i - x_k), 1e-12) over the
// boundary's own range (boundaryVolume.ts:26-52), for the slots of this
// fluid's dynamic boundaries (predicted positions, every substep) or static
// boundaries (committed positions, once). The walk is the reference's
// 27-bucket query with bucket dedup (query.ts:24-70).
template <bool PERSIST>
__global__ __launch_bounds__(PF_BLOCK) void k_boundary_volume(const SceneDev* __restrict__ scenes,
                                                              const Tile* __restrict__ tiles,
                                                              uint32_t ntiles, int fi, int p, int c,
                                                              int dynamic) {
  pf_for_tiles<PERSIST>(ntiles, [&](uint32_t ti) {
    const Tile t = tiles[ti];
    const SceneDev& S = scenes[t.scene];
    const FluidDev& F = S.fluid[fi];
    const uint32_t s = t.first + threadIdx.x;
    if (s >= S.n || F.nboundary == 0u) return;
    const uint32_t id = S.id[p][s];
    int b = -1;
    for (uint32_t k = 0; k < F.nboundary; ++k)
      if (in_range(id, F.bstart[k], F.bcount[k]) && (int)((F.bdynamic >> k) & 1u) == dynamic) b = (int)k;
    if (b < 0) return;
    const uint32_t lo = F.bstart[b], cnt = F.bcount[b];
    const float4* __restrict__ pos = dynamic ? S.xs[c] : S.x[p];
    const float3 xi = pf_xyz(pos[s]);
    const int qx = pf_query_coord(xi.x, S.ox, S.cell);
    const int qy = pf_query_coord(xi.y, S.oy, S.cell);
    const int qz = pf_query_coord(xi.z, S.oz, S.cell);
    uint32_t visited[27];
    uint32_t nvisited = 0;
    float sum = 0.0f;
    for (int dz = -1; dz <= 1; ++dz)
      for (int dy = -1; dy <= 1; ++dy)
        for (int dx = -1; dx <= 1; ++dx) {
          const uint32_t bk = pf_morton((int)((uint32_t)qx + (uint32_t)dx), (int)((uint32_t)qy + (uint32_t)dy),
                                        (int)((uint32_t)qz + (uint32_t)dz)) & S.hash_mask;
          bool seen = false;
          for (uint32_t k = 0; k < nvisited; ++k) seen |= (visited[k] == bk);
          visited[nvisited++] = bk;
          if (seen) continue;
          for (uint32_t j = S.start[bk]; j < S.start[bk + 1]; ++j) {
            if (!in_range(S.id[p][j], lo, cnt)) continue;
            const float3 xj = pf_xyz(pos[j]);
            const float3 d = make_float3(xi.x - xj.x, xi.y - xj.y, xi.z - xj.z);
            const float r2 = pf_dot3(d, d);
            if (r2 >= F.h2) continue;
            sum = sum + pf_poly6(r2, F.h2, F.poly6);
          }
        }
    S.bvol[p][s] = 1.0f / fmaxf(sum, 1e-12f);
  });
}

// Container contact + position accumulator apply (+ advect), in place.
template <bool ADVECT, bool PERSIST>
__global__ __launch_bounds__(PF_BLOCK) void k_walls(const SceneDev* __restrict__ scenes,
                                                    const Tile* __restrict__ tiles, uint32_t ntiles,
                                                    int p, int c, int first_iteration, int substep) {
  pf_for_tiles<PERSIST>(ntiles, [&](uint32_t ti) {
    const Tile t = tiles[ti];
    const SceneDev& S = scenes[t.scene];
    const uint32_t s = t.first + threadIdx.x;
    if (s >= S.n) return;
    const float4 xi4 = S.xs[c][s];
    float3 xn = pf_xyz(xi4);
    const float w = S.w[p][s];
    const float4 x = S.x[p][s];
    int tk[3] = {0, 0, 0};
    if (S.ct.on) {  // the contact solve's sums share the accumulator (loop.ts:183-236)
      const int4 a = S.ct.acc[s];
      S.ct.acc[s] = make_int4(0, 0, 0, 0);
      tk[0] = a.x; tk[1] = a.y; tk[2] = a.z;
    }
    walls_solve(S, s, xn, pf_xyz(x), w, S.clock[substep], first_iteration != 0, tk);
    sdf_solve(S, s, xn, pf_xyz(x), w, first_iteration != 0, tk);
    // A scene in the same launch with neither colliders nor contacts has no
    // accumulator apply in its own schedule.
    if (S.nprim || S.nsdf || S.ct.on) {
      bool sat;
      const float3 ch = ticks_to_change(tk, S.pos_inv, &sat);
      if (sat) flag_set(S, 0, PF_FLAG_POSITIONS);
      xn = pf_add3(xn, ch);
    }
    if (ADVECT && w > 0.0f) {
      S.v[p][s] = make_float4((xn.x - x.x) / S.dt, (xn.y - x.y) / S.dt, (xn.z - x.z) / S.dt, 0.0f);
      S.x[p][s] = make_float4(xn.x, xn.y, xn.z, 0.0f);
    }
    S.xs[c][s] = make_float4(xn.x, xn.y, xn.z, xi4.w);
  });
}

// Advect only (integrate.ts:36-44), for scenes without colliders or fluid.
template <bool PERSIST>
__global__ __launch_bounds__(PF_BLOCK) void k_advect(const SceneDev* __restrict__ scenes,
                                                     const Tile* __restrict__ tiles, uint32_t ntiles,
                                                     int p, int c) {
  pf_for_tiles<PERSIST>(ntiles, [&](uint32_t ti) {
    const Tile t = tiles[ti];
    const SceneDev& S = scenes[t.scene];
    const uint32_t s = t.first + threadIdx.x;
    if (s >= S.n) return;
    const float w = S.w[p][s];
    if (!(w > 0.0f)) return;
    const float4 xsv = S.xs[c][s];
    const float4 x = S.x[p][s];
    S.v[p][s] = make_float4((xsv.x - x.x) / S.dt, (xsv.y - x.y) / S.dt, (xsv.z - x.z) / S.dt, 0.0f);
    S.x[p][s] = make_float4(xsv.x, xsv.y, xsv.z, 0.0f);
  });
}

// ---------------------------------------------------------------------------
// Velocity passes
// ---------------------------------------------------------------------------
// One neighbour's vorticity and XSPH terms (velocity.ts:45-62), added when ok.
template <bool VORT, bool VISC>
static __device__ __forceinline__ void velocity_term(const FluidDev& F, bool ok, bool vort, bool visc,
                                                     float3 xi, float3 vi, float4 xj, float3 vj,
                                                     float vol, float3& curl, float3& sm) {
  const float3 off = make_float3(xi.x - xj.x, xi.y - xj.y, xi.z - xj.z);
  if (VORT && vort)
    curl = sel3(ok, pf_add3(curl, pf_cross3(pf_sub3(vi, vj), pf_mul3(off, pf_spiky_scale(off, F.h, F.spiky) * vol))),
                curl);
  if (VISC && visc)
    sm = sel3(ok, pf_add3(sm, pf_mul3(pf_mul3(pf_sub3(vj, vi), pf_poly6(pf_dot3(off, off), F.h2, F.poly6)), vol)), sm);
}

// BND: some scene of the launch has boundary particles, whose volume differs
// from the fluid's; the union then holds each entry's volume (32 bytes an
// entry), else the volume is the fluid's own and an entry takes 24 bytes,
// which fits three workgroups per CU instead of two.
template <bool VORT, bool VISC, bool BND, bool PERSIST>
__global__ __launch_bounds__(PF_BLOCK) void k_velocity(const SceneDev* __restrict__ scenes,
                                                       const Tile* __restrict__ tiles,
                                                       uint32_t ntiles, int fi, int p, uint32_t cap) {
  extern __shared__ float4 pf_lds[];
  float4* u_x = pf_lds;        // position; the neighbour volume (BND) or velocity x
  float4* u_v = pf_lds + cap;  // velocity (BND)
  float2* u_yz = reinterpret_cast<float2*>(pf_lds + cap);  // velocity y, z (not BND)
  pf_for_tiles<PERSIST>(ntiles, [&](uint32_t ti) {
    const Tile t = tiles[ti];
    const SceneDev& S = scenes[t.scene];
    const FluidDev& F = S.fluid[fi];
    const uint32_t s = t.first + threadIdx.x;
    const uint32_t tile = t.first / PF_BLOCK;
    const uint32_t n = S.n;
    const bool vort = VORT && (F.effects & PF_FX_VORTICITY);
    const bool visc = VISC && (F.effects & PF_FX_VISCOSITY);
    const float4* __restrict__ X = S.x[p];
    const float4* __restrict__ V = S.v[p];
    const uint32_t U = F.tu_size[tile];
    const bool staged = (vort || visc) && U != PF_TU_GLOBAL;
    if (staged) {
      uint32_t js[PF_TU_PER];
      tu_load_slots(F.tu_slots + (size_t)tile * PF_TU_MAX, U, js);
      float4 xv[PF_TU_PER], vv[PF_TU_PER];
      float ov[PF_TU_PER];
  #pragma unroll
      for (int i = 0; i < PF_TU_PER; ++i) {
        xv[i] = X[js[i]];
        vv[i] = V[js[i]];
        ov[i] = F.pv;
      }
      if (BND && S.has_boundary) {
  #pragma unroll
        for (int i = 0; i < PF_TU_PER; ++i) ov[i] = neighbor_volume(S, F, p, js[i]);
      }
  #pragma unroll
      for (int i = 0; i < PF_TU_PER; ++i) {
        const uint32_t u = threadIdx.x + i * PF_BLOCK;
        if (u < U) {
          if (BND) {
            u_x[u] = make_float4(xv[i].x, xv[i].y, xv[i].z, ov[i]);
            u_v[u] = pf_f4(vv[i]);
          } else {
            u_x[u] = make_float4(xv[i].x, xv[i].y, xv[i].z, vv[i].x);
            u_yz[u] = make_float2(vv[i].y, vv[i].z);
          }
        }
      }
      __syncthreads();
    }
    if (s >= n) return;
    if (!in_range(S.id[p][s], F.start, F.count)) {
      // |omega| of particles outside the fluid reads as zero in the gradient
      // pass (the reference's per-fluid buffer is zero there); slots change
      // particles every substep, so clear it.
      if (vort) F.omega[s] = make_float4(0.f, 0.f, 0.f, 0.f);
      return;
    }
    if (!(vort || visc)) return;
    const float3 xi = pf_xyz(X[s]);
    const float3 vi = pf_xyz(V[s]);
    const uint32_t cnt = F.nl_count[s];
    float3 curl = make_float3(0.f, 0.f, 0.f);
    float3 sm = make_float3(0.f, 0.f, 0.f);
    if (staged) {
      Nl16Cursor cur(F.nl16, n, s, U);
      // Whole groups of four need no per-entry test; the last partial group does.
      auto group = [&](uint32_t k0, bool tail) {
        uint32_t idx[4];
        cur.next(k0, idx);
        float4 xj[4];
        float3 vj[4];
        float vol[4];
  #pragma unroll
        for (int u = 0; u < 4; ++u) {
          xj[u] = u_x[idx[u]];
          if (BND) {
            const float4 v4 = u_v[idx[u]];
            vj[u] = pf_xyz(v4);
            vol[u] = xj[u].w;
          } else {
            const float2 yz = u_yz[idx[u]];
            vj[u] = make_float3(xj[u].w, yz.x, yz.y);
            vol[u] = F.pv;
          }
        }
  #pragma unroll
        for (int u = 0; u < 4; ++u)
          velocity_term<VORT, VISC>(F, !tail || k0 + u < cnt, vort, visc, xi, vi, xj[u], vj[u], vol[u], curl, sm);
      };
      const uint32_t full = cnt & ~3u;
      for (uint32_t k0 = 0; k0 < full; k0 += 4) group(k0, false);
      if (full < cnt) group(full, true);
    } else {
      NlCursor cur(F.nl, n, s, cnt);
      for (uint32_t k0 = 0; k0 < cnt; k0 += 4) {
        uint32_t jj[4];
        cur.next(k0, jj);
        float4 xj[4], vj4[4];
        float vol[4];
  #pragma unroll
        for (int u = 0; u < 4; ++u) {
          xj[u] = X[jj[u]];
          vj4[u] = V[jj[u]];
          vol[u] = neighbor_volume(S, F, p, jj[u]);
        }
  #pragma unroll
        for (int u = 0; u < 4; ++u)
          velocity_term<VORT, VISC>(F, k0 + u < cnt, vort, visc, xi, vi, xj[u], pf_xyz(vj4[u]), vol[u], curl, sm);
      }
    }
    if (vort) F.omega[s] = make_float4(curl.x, curl.y, curl.z, pf_len3(curl));
    if (visc) F.dv[s] = make_float4(sm.x * F.visc, sm.y * F.visc, sm.z * F.visc, 0.0f);
  });
}

// One neighbour's term of the vorticity gradient (velocity.ts:80-88), added when ok.
static __device__ __forceinline__ void eta_term(const FluidDev& F, bool ok, float3 xi, float oi,
                                                float4 xj, float oj, float3& eta) {
  const float3 d = make_float3(xi.x - xj.x, xi.y - xj.y, xi.z - xj.z);
  eta = sel3(ok, pf_add3(eta, pf_mul3(d, pf_spiky_scale(d, F.h, F.spiky) * ((oj - oi) * F.pv))), eta);
}

template <bool VORT, bool VISC, bool FRICTION, bool PERSIST>
__global__ __launch_bounds__(PF_BLOCK) void k_vorticity(const SceneDev* __restrict__ scenes,
                                                        const Tile* __restrict__ tiles,
                                                        uint32_t ntiles, int fi, int p, int c,
                                                        int substep, int touch_valid, uint32_t cap) {
  extern __shared__ float4 pf_lds[];
  float4* u_x = pf_lds;                                       // position, |omega|
  uint8_t* u_skip = reinterpret_cast<uint8_t*>(pf_lds + cap);  // boundary particle (left out of the gradient)
  pf_for_tiles<PERSIST>(ntiles, [&](uint32_t ti) {
    const Tile t = tiles[ti];
    const SceneDev& S = scenes[t.scene];
    const FluidDev& F = S.fluid[fi];
    const uint32_t s = t.first + threadIdx.x;
    const uint32_t tile = t.first / PF_BLOCK;
    const uint32_t n = S.n;
    const bool vort = VORT && (F.effects & PF_FX_VORTICITY);
    const bool visc = VISC && (F.effects & PF_FX_VISCOSITY);
    const float4* __restrict__ X = S.x[p];
    const uint32_t U = F.tu_size[tile];
    const bool staged = vort && U != PF_TU_GLOBAL;
    if (staged) {
      uint32_t js[PF_TU_PER];
      tu_load_slots(F.tu_slots + (size_t)tile * PF_TU_MAX, U, js);
      float4 xv[PF_TU_PER];
      float ov[PF_TU_PER];
      uint8_t kv[PF_TU_PER];
  #pragma unroll
      for (int i = 0; i < PF_TU_PER; ++i) {
        xv[i] = X[js[i]];
        ov[i] = F.omega[js[i]].w;
        kv[i] = 0;
      }
      if (S.has_boundary) {
  #pragma unroll
        for (int i = 0; i < PF_TU_PER; ++i) kv[i] = S.bvol[p][js[i]] > 0.0f ? 1 : 0;
      }
  #pragma unroll
      for (int i = 0; i < PF_TU_PER; ++i) {
        const uint32_t u = threadIdx.x + i * PF_BLOCK;
        if (u < U) {
          u_x[u] = make_float4(xv[i].x, xv[i].y, xv[i].z, ov[i]);
          u_skip[u] = kv[i];
        }
      }
      __syncthreads();
    }
    if (s >= n) return;
    float4 v = S.v[p][s];
    if ((vort || visc) && in_range(S.id[p][s], F.start, F.count)) {
      float3 change = make_float3(0.f, 0.f, 0.f);
      if (vort) {
        const float3 xi = pf_xyz(X[s]);
        const float4 oi4 = F.omega[s];
        const float oi = oi4.w;
        const uint32_t cnt = F.nl_count[s];
        float3 eta = make_float3(0.f, 0.f, 0.f);
        if (staged) {
          Nl16Cursor cur(F.nl16, n, s, U);
          // Whole groups of four in a scene without boundary particles need no
          // per-entry test; the last partial group, or skipped boundary
          // neighbours, do.
          auto group = [&](uint32_t k0, bool tail, bool skips) {
            uint32_t idx[4];
            cur.next(k0, idx);
            float4 xj[4];
            bool ok[4];
  #pragma unroll
            for (int u = 0; u < 4; ++u) {
              xj[u] = u_x[idx[u]];
              ok[u] = (!tail || k0 + u < cnt) && !(skips && u_skip[idx[u]]);
            }
  #pragma unroll
            for (int u = 0; u < 4; ++u) eta_term(F, ok[u], xi, oi, xj[u], xj[u].w, eta);
          };
          const uint32_t full = cnt & ~3u;
          if (S.has_boundary) {
            for (uint32_t k0 = 0; k0 < full; k0 += 4) group(k0, false, true);
          } else {
            for (uint32_t k0 = 0; k0 < full; k0 += 4) group(k0, false, false);
          }
          if (full < cnt) group(full, true, S.has_boundary != 0);
        } else {
          NlCursor cur(F.nl, n, s, cnt);
          for (uint32_t k0 = 0; k0 < cnt; k0 += 4) {
            uint32_t jj[4];
            cur.next(k0, jj);
            float4 xj[4];
            float oj[4];
            bool ok[4];
  #pragma unroll
            for (int u = 0; u < 4; ++u) {
              xj[u] = X[jj[u]];
              oj[u] = F.omega[jj[u]].w;
              ok[u] = (k0 + u < cnt) && !(S.has_boundary && S.bvol[p][jj[u]] > 0.0f);
            }
  #pragma unroll
            for (int u = 0; u < 4; ++u) eta_term(F, ok[u], xi, oi, xj[u], oj[u], eta);
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
      sdf_friction(S, s, pf_xyz(S.xs[c][s]), pf_xyz(v), S.w[p][s], tk);
      bool sat;
      const float3 ch = ticks_to_change(tk, S.vel_inv, &sat);
      if (sat) flag_set(S, 0, PF_FLAG_VELOCITIES);
      v.x = v.x + ch.x;
      v.y = v.y + ch.y;
      v.z = v.z + ch.z;
    }
    S.v[p][s] = v;
  });
}

// Friction + velocity accumulator apply on its own (no fluid to fuse with).
template <bool PERSIST>
__global__ __launch_bounds__(PF_BLOCK) void k_friction(const SceneDev* __restrict__ scenes,
                                                       const Tile* __restrict__ tiles,
                                                       uint32_t ntiles, int p, int c, int substep) {
  pf_for_tiles<PERSIST>(ntiles, [&](uint32_t ti) {
    const Tile t = tiles[ti];
    const SceneDev& S = scenes[t.scene];
    const uint32_t s = t.first + threadIdx.x;
    if (s >= S.n) return;
    if (!(S.nprim || S.nsdf || S.ct.on)) return;  // nothing to apply in this scene
    float4 v = S.v[p][s];
    int tk[3] = {0, 0, 0};
    if (S.ct.on) {  // the contacts' kinetic friction sums (loop.ts:226-230)
      const int4 a = S.ct.acc[s];
      S.ct.acc[s] = make_int4(0, 0, 0, 0);
      tk[0] = a.x; tk[1] = a.y; tk[2] = a.z;
    }
    walls_friction(S, s, pf_xyz(S.xs[c][s]), pf_xyz(v), S.w[p][s], S.clock[substep], tk);
    sdf_friction(S, s, pf_xyz(S.xs[c][s]), pf_xyz(v), S.w[p][s], tk);
    bool sat;
    const float3 ch = ticks_to_change(tk, S.vel_inv, &sat);
    if (sat) flag_set(S, 0, PF_FLAG_VELOCITIES);
    v.x = v.x + ch.x;
    v.y = v.y + ch.y;
    v.z = v.z + ch.z;
    S.v[p][s] = v;
  });
}

// ---------------------------------------------------------------------------
// Host launchers
// ---------------------------------------------------------------------------
// A tile kernel K (with its leading template arguments TARGS, parenthesised)
// over the tiles, in the launch shape's form (launch_tiles).
#define PF_UNPAREN(...) __VA_ARGS__
#define PF_TILES(K, TARGS, lds, ...) \
  launch_tiles(K<PF_UNPAREN TARGS false>, K<PF_UNPAREN TARGS true>, ntiles, (lds), st, __VA_ARGS__)

void launch_neighbors(hipStream_t st, const SceneDev* sc, const Tile* tl, uint32_t ntiles, int fi,
                      int p, int c, bool normal, uint32_t cap, int snap) {
  if (!ntiles) return;
  // x, y, z of cap union entries; at least the leader test's 4 words per thread
  const size_t lds = std::max<size_t>((size_t)cap * 3 * sizeof(float), 4 * PF_BLOCK * sizeof(uint32_t));
  if (normal) PF_TILES(k_neighbors, (true,), lds, sc, tl, ntiles, fi, p, c, cap, snap);
  else PF_TILES(k_neighbors, (false,), lds, sc, tl, ntiles, fi, p, c, cap, snap);
}

void launch_impulse_apply(hipStream_t st, const SceneDev* sc, const Tile* tl, uint32_t ntiles, int fi, int p,
                          int c) {
  if (!ntiles) return;
  PF_TILES(k_impulse_apply, (), 0, sc, tl, ntiles, fi, p, c);
}

void launch_impulses(hipStream_t st, const SceneDev* sc, const Tile* tl, uint32_t ntiles, int fi,
                     int p, int c, bool surface_tension, bool adhesion) {
  if (!ntiles) return;
#define PF_IK(a, b) PF_TILES(k_impulses, (a, b,), 0, sc, tl, ntiles, fi, p, c)
  if (surface_tension && adhesion) PF_IK(true, true);
  else if (surface_tension) PF_IK(true, false);
  else if (adhesion) PF_IK(false, true);
  else PF_IK(false, false);
#undef PF_IK
}

void launch_lambda(hipStream_t st, const SceneDev* sc, const Tile* tl, uint32_t ntiles, int fi,
                   int p, int c, uint32_t cap) {
  if (!ntiles) return;
  const size_t lds = (size_t)cap * (sizeof(float4) + sizeof(float));
  PF_TILES(k_lambda, (), lds, sc, tl, ntiles, fi, p, c, cap);
}

void launch_delta(hipStream_t st, const SceneDev* sc, const Tile* tl, uint32_t ntiles, int fi,
                  int p, int c, bool walls, bool advect, int first_iteration, int substep, uint32_t cap) {
  if (!ntiles) return;
  const size_t lds = (size_t)cap * (sizeof(float4) + sizeof(float));
#define PF_DK(w, a) PF_TILES(k_delta, (w, a,), lds, sc, tl, ntiles, fi, p, c, first_iteration, substep, cap)
  if (walls && advect) PF_DK(true, true);
  else if (walls) PF_DK(true, false);
  else if (advect) PF_DK(false, true);
  else PF_DK(false, false);
#undef PF_DK
}

void launch_walls(hipStream_t st, const SceneDev* sc, const Tile* tl, uint32_t ntiles, int p, int c,
                  bool advect, int first_iteration, int substep) {
  if (!ntiles) return;
  if (advect) PF_TILES(k_walls, (true,), 0, sc, tl, ntiles, p, c, first_iteration, substep);
  else PF_TILES(k_walls, (false,), 0, sc, tl, ntiles, p, c, first_iteration, substep);
}

void launch_reaction_apply(hipStream_t st, const SceneDev* sc, const Tile* tl, uint32_t ntiles, int fi, int p,
                           int c) {
  if (!ntiles) return;
  PF_TILES(k_reaction_apply, (), 0, sc, tl, ntiles, fi, p, c);
}

void launch_boundary_volume(hipStream_t st, const SceneDev* sc, const Tile* tl, uint32_t ntiles, int fi, int p,
                            int c, int dynamic) {
  if (!ntiles) return;
  PF_TILES(k_boundary_volume, (), 0, sc, tl, ntiles, fi, p, c, dynamic);
}

void launch_advect(hipStream_t st, const SceneDev* sc, const Tile* tl, uint32_t ntiles, int p, int c) {
  if (!ntiles) return;
  PF_TILES(k_advect, (), 0, sc, tl, ntiles, p, c);
}

void launch_velocity(hipStream_t st, const SceneDev* sc, const Tile* tl, uint32_t ntiles, int fi,
                     int p, bool vort, bool visc, bool boundary, uint32_t cap) {
  if (!ntiles || !(vort || visc)) return;
  const size_t lds = (size_t)cap * (boundary ? 2 * sizeof(float4) : sizeof(float4) + sizeof(float2));
#define PF_VK(a, b, c) PF_TILES(k_velocity, (a, b, c,), lds, sc, tl, ntiles, fi, p, cap)
  if (boundary) {
    if (vort && visc) PF_VK(true, true, true);
    else if (vort) PF_VK(true, false, true);
    else PF_VK(false, true, true);
  } else {
    if (vort && visc) PF_VK(true, true, false);
    else if (vort) PF_VK(true, false, false);
    else PF_VK(false, true, false);
  }
#undef PF_VK
}

void launch_vorticity(hipStream_t st, const SceneDev* sc, const Tile* tl, uint32_t ntiles, int fi,
                      int p, int c, bool vort, bool visc, bool friction, int substep,
                      int touch_valid, uint32_t cap) {
  if (!ntiles) return;
  const size_t lds = (size_t)cap * sizeof(float4) + ((cap + 15u) & ~15u);
#define PF_VK(a, b, f) PF_TILES(k_vorticity, (a, b, f,), lds, sc, tl, ntiles, fi, p, c, substep, touch_valid, cap)
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

void launch_viscosity_solve(hipStream_t st, const SceneDev* sc, const Tile* tl, uint32_t ntiles, int fi,
                            int p, uint32_t sweeps) {
  if (!ntiles || !sweeps) return;
  PF_TILES(k_isv_seed, (), 0, sc, tl, ntiles, fi, p);
  for (uint32_t k = 0; k < sweeps; ++k) PF_TILES(k_isv_sweep, (), 0, sc, tl, ntiles, fi, p, k);
  PF_TILES(k_isv_apply, (), 0, sc, tl, ntiles, fi, p, sweeps);
}

void launch_friction(hipStream_t st, const SceneDev* sc, const Tile* tl, uint32_t ntiles, int p,
                     int c, int substep) {
  if (!ntiles) return;
  PF_TILES(k_friction, (), 0, sc, tl, ntiles, p, c, substep);
}

}  // namespace pf
