// Hash-grid counting sort for gfx950.
//
// Replaces the reference's 8-dispatch rebuild (HashGrid.ts:195-204:
// resetCounts, resetOverflow, cellIndexAndHistogram, blockScan, blockSumScan,
// finalizeCellRanges, scatter, sortedPositions) with four kernels:
//
//   k_hash     predict (scene mode) + Morton bucket + wave-aggregated histogram;
//              the histogram atomic's return value is the arrival rank.
//   k_scan     single-pass decoupled look-back exclusive scan of the counts
//              (no 2^20 bucket cap, sort.ts:22-28), also clears the counts.
//   k_place    bucket position = start + arrival rank.
//   k_rank     stable fix-up: a particle's final slot inside its bucket is the
//              number of bucket members with a smaller original id, so the
//              permutation is canonical (ascending id within a bucket), then
//              the particle state is gathered into the new cell-sorted order.
//
// Outputs match the reference bit-for-bit for keys, counts, offsets and the
// sorted key sequence; the permutation is the stable member of the set the
// reference's atomic scatter may produce (sort.ts:187-192).
#include "pf_internal.hpp"
#include "pf_kernels.hpp"

namespace pf {

static __device__ __forceinline__ uint32_t* flag_word(const SceneDev& S, int parity, int f) {
  return S.flags + parity * PF_NUM_FLAGS + f;
}

// Bucket of a position: cellIndex.ts:113-143.
static __device__ __forceinline__ uint32_t bucket_of(const SceneDev& S, float px, float py, float pz,
                                                     uint32_t* over) {
  const int cx = pf_bin_coord(px, S.ox, S.cell, over);
  const int cy = pf_bin_coord(py, S.oy, S.cell, over);
  const int cz = pf_bin_coord(pz, S.oz, S.cell, over);
  return pf_morton(cx, cy, cz) & S.hash_mask;
}

// Predicted position of a slot (integrate.ts:25-34). Pinned particles keep
// their stored prediction.
static __device__ __forceinline__ void predict(const SceneDev& S, float4 x, float4 v, float w,
                                               float4 xs_prev, float4* vel_out, float4* xs_out) {
  if (w > 0.0f) {
    const float vx = v.x + S.gx * S.dt;
    const float vy = v.y + S.gy * S.dt;
    const float vz = v.z + S.gz * S.dt;
    *vel_out = make_float4(vx, vy, vz, 0.0f);
    *xs_out = make_float4(x.x + vx * S.dt, x.y + vy * S.dt, x.z + vz * S.dt, 0.0f);
  } else {
    *vel_out = v;
    *xs_out = make_float4(xs_prev.x, xs_prev.y, xs_prev.z, 0.0f);
  }
}

// MODE 0: standalone grid over x[0] positions (ids = slots).
// MODE 1: scene substep; predicts first, reads state buffers of `parity`.
template <int MODE, bool PERSIST>
__global__ __launch_bounds__(PF_BLOCK) void k_hash(const SceneDev* __restrict__ scenes,
                                                   const Tile* __restrict__ tiles, uint32_t ntiles,
                                                   int parity, int xs_prev, int reset_step_flags) {
  pf_for_tiles<PERSIST>(ntiles, [&](uint32_t ti) {
    const Tile t = tiles[ti];
    const SceneDev& S = scenes[t.scene];
    const uint32_t i = t.first + threadIdx.x;
    const bool valid = i < S.n;
    uint32_t over = 0;
    uint32_t b = 0xffffffffu;
    if (MODE == 1 && reset_step_flags && t.first == 0 && threadIdx.x == 0) {
      // Step start: position/velocity saturation flags cover a whole step (loop.ts:206-210).
      __hip_atomic_store(flag_word(S, 0, PF_FLAG_POSITIONS), 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_store(flag_word(S, 0, PF_FLAG_VELOCITIES), 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    if (MODE == 1 && S.ct.on && t.first == 0 && threadIdx.x == 0) {
      // Substep start: the contact counter (resetCounterKernel, ContactBuffer.ts:62-65).
      __hip_atomic_store(S.ct.count, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    if (valid) {
      float4 p;
      if (MODE == 0) {
        p = S.x[0][i];
      } else {
        const float w = S.w[parity][i];
        float4 vel;
        const float4 xsp = (w > 0.0f) ? make_float4(0.f, 0.f, 0.f, 0.f) : S.xs[xs_prev][i];
        predict(S, S.x[parity][i], S.v[parity][i], w, xsp, &vel, &p);
      }
      b = bucket_of(S, p.x, p.y, p.z, &over);
      S.key[i] = b;
    }
    // Wave-level aggregation: runs of equal keys in adjacent lanes (the common
    // case once particles are cell-sorted) issue one atomic per run.
    const uint32_t lane = __lane_id();
    const uint32_t prevb = __shfl_up(b, 1);
    const bool head = valid && (lane == 0 || prevb != b);
    const unsigned long long heads = __ballot(head);
    const unsigned long long validm = __ballot(valid);
    const uint32_t nvalid = __popcll(validm);
    const unsigned long long below = (lane == 63) ? ~0ull : ((2ull << lane) - 1ull);
    const uint32_t head_lane = valid ? (63u - __clzll(heads & below)) : 0u;
    uint32_t base = 0;
    if (head) {
      const unsigned long long above = heads & ~below;
      const uint32_t run_end = above ? (uint32_t)__ffsll((long long)above) - 1u : nvalid;
      base = atomicAdd(&S.counts[b], run_end - lane);
    }
    const uint32_t hb = __shfl(base, (int)head_lane);
    if (valid) S.rank[i] = hb + (lane - head_lane);
    if (over) pf_flag_raise(flag_word(S, MODE == 1 ? parity : 0, PF_FLAG_GRID));
  });
}

// ---------------------------------------------------------------------------
// Decoupled look-back scan (single pass). Status word per tile:
//   bits 0..31 value, bits 32..33 flag (1 aggregate, 2 inclusive), 34.. epoch.
// Tile order comes from an atomic ticket so every tile's predecessors are
// already running. Each scene owns an epoch that its last finishing tile
// advances, and the launch's last tile resets the ticket, so nothing needs a
// reset between launches (also under graph replay), and scenes scanned by
// different launches never see each other's words.
// ---------------------------------------------------------------------------
struct __align__(16) ScanCtlDev {
  uint32_t ticket, done, pad0, pad1;
};

static __device__ __forceinline__ unsigned long long pack_status(uint32_t epoch, uint32_t flag,
                                                                 uint32_t value) {
  return ((unsigned long long)(epoch & 0x3fffffffu) << 34) | ((unsigned long long)flag << 32) |
         (unsigned long long)value;
}

__global__ __launch_bounds__(PF_BLOCK) void k_scan(const SceneDev* __restrict__ scenes,
                                                   const uint2* __restrict__ stiles,
                                                   uint32_t nstiles, ScanCtlDev* ctl) {
  __shared__ uint32_t s_ticket, s_epoch, s_prefix;
  __shared__ uint32_t s_wave[PF_BLOCK / 64];
  if (threadIdx.x == 0) {
    const uint32_t tk = atomicAdd(&ctl->ticket, 1u);
    s_ticket = tk;
    s_epoch = __hip_atomic_load(&scenes[stiles[tk].x].scan_ctl[1], __ATOMIC_RELAXED,
                                __HIP_MEMORY_SCOPE_AGENT);
  }
  __syncthreads();
  const uint32_t ticket = s_ticket;
  const uint32_t epoch = s_epoch;
  const uint2 st = stiles[ticket];
  const SceneDev& S = scenes[st.x];
  const uint32_t H = S.H;
  const uint32_t tile = st.y;
  const uint32_t base = tile * PF_SCAN_TILE + threadIdx.x * 16u;

  uint32_t c[16];
  if (base + 16u <= H) {
    const uint4* src = reinterpret_cast<const uint4*>(S.counts + base);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const uint4 q = src[k];
      c[4 * k + 0] = q.x; c[4 * k + 1] = q.y; c[4 * k + 2] = q.z; c[4 * k + 3] = q.w;
    }
    uint4* dst = reinterpret_cast<uint4*>(S.counts + base);
#pragma unroll
    for (int k = 0; k < 4; ++k) dst[k] = make_uint4(0u, 0u, 0u, 0u);
  } else {
#pragma unroll
    for (int k = 0; k < 16; ++k) {
      const uint32_t idx = base + (uint32_t)k;
      c[k] = idx < H ? S.counts[idx] : 0u;
      if (idx < H) S.counts[idx] = 0u;
    }
  }
  uint32_t tsum = 0;
#pragma unroll
  for (int k = 0; k < 16; ++k) tsum += c[k];

  // Block-wide exclusive scan of the per-thread sums (wave64 shuffles + LDS).
  const uint32_t lane = __lane_id();
  const uint32_t wave = threadIdx.x >> 6;
  uint32_t incl = tsum;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const uint32_t y = __shfl_up(incl, d);
    if (lane >= (uint32_t)d) incl += y;
  }
  if (lane == 63) s_wave[wave] = incl;
  __syncthreads();
  uint32_t wave_off = 0, aggregate = 0;
#pragma unroll
  for (int w = 0; w < PF_BLOCK / 64; ++w) {
    const uint32_t tot = s_wave[w];
    if ((uint32_t)w < wave) wave_off += tot;
    aggregate += tot;
  }
  const uint32_t thread_excl = wave_off + incl - tsum;

  unsigned long long* status = S.scan_status;
  if (wave == 0) {
    uint32_t prefix = 0;
    if (tile == 0) {
      if (lane == 0)
        __hip_atomic_store(&status[0], pack_status(epoch, 2u, aggregate), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
    } else {
      if (lane == 0)
        __hip_atomic_store(&status[tile], pack_status(epoch, 1u, aggregate), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
      int pos = (int)tile - 1;
      while (true) {
        const int idx = pos - (int)lane;
        uint32_t flag = 2u, value = 0u;
        if (idx >= 0) {
          const unsigned long long s =
              __hip_atomic_load(&status[idx], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          const uint32_t e = (uint32_t)(s >> 34);
          flag = (e == (epoch & 0x3fffffffu)) ? (uint32_t)((s >> 32) & 3u) : 0u;
          value = (uint32_t)s;
        }
        const unsigned long long incl_m = __ballot(flag == 2u);
        const unsigned long long wait_m = __ballot(flag == 0u);
        const uint32_t first_incl = incl_m ? (uint32_t)__ffsll((long long)incl_m) - 1u : 64u;
        const unsigned long long upto =
            first_incl >= 63u ? ~0ull : ((2ull << first_incl) - 1ull);
        if (wait_m & upto) continue;  // a predecessor has not published yet
        uint32_t v = (lane <= first_incl) ? value : 0u;
#pragma unroll
        for (int d = 32; d > 0; d >>= 1) v += __shfl_xor(v, d);
        prefix += v;
        if (incl_m) break;
        pos -= 64;
      }
      if (lane == 0)
        __hip_atomic_store(&status[tile], pack_status(epoch, 2u, prefix + aggregate),
                           __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    if (lane == 0) s_prefix = prefix;
  }
  __syncthreads();
  uint32_t run = s_prefix + thread_excl;
  if (base + 16u <= H) {
    uint32_t o[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) { o[k] = run; run += c[k]; }
    uint4* dst = reinterpret_cast<uint4*>(S.start + base);
#pragma unroll
    for (int k = 0; k < 4; ++k) dst[k] = make_uint4(o[4 * k], o[4 * k + 1], o[4 * k + 2], o[4 * k + 3]);
  } else {
#pragma unroll
    for (int k = 0; k < 16; ++k) {
      const uint32_t idx = base + (uint32_t)k;
      if (idx < H) S.start[idx] = run;
      run += c[k];
    }
  }
  if (tile == S.nscan_tiles - 1 && threadIdx.x == 0) S.start[H] = s_prefix + aggregate;
  if (threadIdx.x == 0) {
    const uint32_t ds = atomicAdd(&S.scan_ctl[0], 1u);
    if (ds == S.nscan_tiles - 1u) {
      atomicExch(&S.scan_ctl[0], 0u);
      atomicAdd(&S.scan_ctl[1], 1u);
    }
    const uint32_t d = atomicAdd(&ctl->done, 1u);
    if (d == nstiles - 1u) {
      atomicExch(&ctl->ticket, 0u);
      atomicExch(&ctl->done, 0u);
    }
  }
}

// Bucket position from start + arrival rank (the reference's atomic scatter).
template <int MODE, bool PERSIST>
__global__ __launch_bounds__(PF_BLOCK) void k_place(const SceneDev* __restrict__ scenes,
                                                    const Tile* __restrict__ tiles, uint32_t ntiles,
                                                    int parity) {
  pf_for_tiles<PERSIST>(ntiles, [&](uint32_t ti) {
    const Tile t = tiles[ti];
    const SceneDev& S = scenes[t.scene];
    const uint32_t i = t.first + threadIdx.x;
    if (i >= S.n) return;
    const uint32_t b = S.key[i];
    const uint32_t slot = S.start[b] + S.rank[i];
    S.tmp_src[slot] = i;
    S.tmp_key[slot] = b;
    if (MODE == 1) S.tmp_id[slot] = S.id[parity][i];
  });
}

// Stable rank inside the bucket, then gather into cell-sorted order.
// MODE 0 writes the permutation / sorted keys (/ sorted positions).
// MODE 1 gathers the particle state of `parity` into `parity ^ 1`, predicting
// velocities and positions (integrate.ts:25-34) on the way.
template <int MODE, bool PERSIST>
__global__ __launch_bounds__(PF_BLOCK) void k_rank(const SceneDev* __restrict__ scenes,
                                                   const Tile* __restrict__ tiles, uint32_t ntiles,
                                                   int parity, int xs_prev, int xs_out,
                                                   uint32_t* __restrict__ out_perm,
                                                   uint32_t* __restrict__ out_skey,
                                                   float4* __restrict__ out_spos) {
  pf_for_tiles<PERSIST>(ntiles, [&](uint32_t ti) {
    const Tile t = tiles[ti];
    const SceneDev& S = scenes[t.scene];
    if (MODE == 1 && t.first == 0 && threadIdx.x == 0) {
      // Per-rebuild flags for the next substep start cleared (HashGrid.ts:195-204,
      // neighborList.ts:47): the next substep writes the other parity's words.
      uint32_t* nf = S.flags + (parity ^ 1) * PF_NUM_FLAGS;
      nf[PF_FLAG_GRID] = 0u;
      nf[PF_FLAG_NEIGHBORS] = 0u;
      nf[PF_FLAG_REACTION] = 0u;
      nf[PF_FLAG_LONG_BUCKET] = 0u;
    }
    const uint32_t s = t.first + threadIdx.x;
    if (s >= S.n) return;
    const uint32_t b = S.tmp_key[s];
    const uint32_t* cmp = (MODE == 0) ? S.tmp_src : S.tmp_id;
    const uint32_t my = cmp[s];
    const uint32_t src = S.tmp_src[s];
    // The particle's state is gathered while the rank is counted.
    float4 x, v, xsp;
    float w = 0.0f;
    if (MODE == 1) {
      x = S.x[parity][src];
      v = S.v[parity][src];
      w = S.w[parity][src];
      xsp = S.xs[xs_prev][src];
    }
    const uint32_t st = S.start[b];
    const uint32_t en = S.start[b + 1];
    uint32_t dest;
    if (en - st <= (uint32_t)PF_STABLE_SEGMENT_LIMIT) {
      // Eight of the bucket's keys in flight per round trip.
      const PF_G uint32_t* __restrict__ g_cmp = pf_g(cmp);
      uint32_t r = 0;
      for (uint32_t q0 = st; q0 < en; q0 += 8u) {
        uint32_t k[8];
#pragma unroll
        for (uint32_t u = 0; u < 8; ++u) k[u] = g_cmp[min(q0 + u, en - 1u)];
#pragma unroll
        for (uint32_t u = 0; u < 8; ++u) r += (q0 + u < en && k[u] < my) ? 1u : 0u;
      }
      dest = st + r;
    } else {
      dest = s;
      pf_flag_raise(S.flags + (MODE == 1 ? parity : 0) * PF_NUM_FLAGS + PF_FLAG_LONG_BUCKET);
    }
    if (MODE == 0) {
      out_perm[dest] = src;
      out_skey[s] = b;
      if (out_spos) out_spos[dest] = S.x[0][src];
    } else {
      const int q = parity ^ 1;
      if (w > 0.0f) xsp = make_float4(0.f, 0.f, 0.f, 0.f);
      float4 vel, xs;
      predict(S, x, v, w, xsp, &vel, &xs);
      S.x[q][dest] = x;
      S.v[q][dest] = vel;
      S.w[q][dest] = w;
      S.mass[dest] = (w > 0.0f) ? (1.0f / w) : 0.0f;  // shared.ts:59
      S.id[q][dest] = my;
      S.slot_of[my] = dest;
      S.rho[q][dest] = S.rho[parity][src];
      S.bvol[q][dest] = S.bvol[parity][src];
      S.grp[q][dest] = S.grp[parity][src];
      S.xs[xs_out][dest] = xs;
    }
  });
}

// ---------------------------------------------------------------------------
// Host launchers
// ---------------------------------------------------------------------------
void launch_hash(hipStream_t st, const SceneDev* scenes, const Tile* tiles, uint32_t ntiles,
                 int mode, int parity, int xs_prev, int reset_step_flags) {
  if (ntiles == 0) return;
  if (mode == 0)
    launch_tiles(k_hash<0, false>, k_hash<0, true>, ntiles, 0, st, scenes, tiles, ntiles, parity,
                 xs_prev, reset_step_flags);
  else
    launch_tiles(k_hash<1, false>, k_hash<1, true>, ntiles, 0, st, scenes, tiles, ntiles, parity,
                 xs_prev, reset_step_flags);
}

void launch_scan(hipStream_t st, const SceneDev* scenes, const uint2* stiles, uint32_t nstiles,
                 void* ctl) {
  if (nstiles == 0) return;
  hipLaunchKernelGGL(k_scan, dim3(nstiles), dim3(PF_BLOCK), 0, st, scenes, stiles, nstiles,
                     reinterpret_cast<ScanCtlDev*>(ctl));
}

void launch_place(hipStream_t st, const SceneDev* scenes, const Tile* tiles, uint32_t ntiles,
                  int mode, int parity) {
  if (ntiles == 0) return;
  if (mode == 0)
    launch_tiles(k_place<0, false>, k_place<0, true>, ntiles, 0, st, scenes, tiles, ntiles, parity);
  else
    launch_tiles(k_place<1, false>, k_place<1, true>, ntiles, 0, st, scenes, tiles, ntiles, parity);
}

void launch_rank(hipStream_t st, const SceneDev* scenes, const Tile* tiles, uint32_t ntiles,
                 int mode, int parity, int xs_prev, int xs_out, uint32_t* out_perm,
                 uint32_t* out_skey, float4* out_spos) {
  if (ntiles == 0) return;
  if (mode == 0)
    launch_tiles(k_rank<0, false>, k_rank<0, true>, ntiles, 0, st, scenes, tiles, ntiles, parity,
                 xs_prev, xs_out, out_perm, out_skey, out_spos);
  else
    launch_tiles(k_rank<1, false>, k_rank<1, true>, ntiles, 0, st, scenes, tiles, ntiles, parity,
                 xs_prev, xs_out, out_perm, out_skey, out_spos);
}

}  // namespace pf
