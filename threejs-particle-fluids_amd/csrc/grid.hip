// Hash-grid sorts for gfx950.
//
// Replace the reference's 8-dispatch rebuild (HashGrid.ts:195-204:
// resetCounts, resetOverflow, cellIndexAndHistogram, blockScan, blockSumScan,
// finalizeCellRanges, scatter, sortedPositions) in two forms.
//
// A scene's substep sorts its particles, nearly in cell order already from
// the previous substep, with a counting sort of four kernels:
//
//   k_hash     predict + Morton bucket + wave-aggregated histogram; the
//              histogram atomic's return value is the arrival rank.
//   k_scan     single-pass decoupled look-back exclusive scan of the counts
//              (no 2^20 bucket cap, sort.ts:22-28), also clears the counts.
//   k_place    bucket position = start + arrival rank.
//   k_rank     stable fix-up: a particle's final slot inside its bucket is the
//              number of bucket members with a smaller original id, so the
//              permutation is canonical (ascending id within a bucket), then
//              the particle state is gathered into the new cell-sorted order.
//
// A standalone grid (HashGrid.rebuild over points in any order) runs an LSD
// radix sort instead (k_rs_*, below).
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

// Predicts first, reads the state buffers of `parity`.
template <bool PERSIST>
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
    if (reset_step_flags && t.first == 0 && threadIdx.x == 0) {
      // Step start: position/velocity saturation flags cover a whole step (loop.ts:206-210).
      __hip_atomic_store(flag_word(S, 0, PF_FLAG_POSITIONS), 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_store(flag_word(S, 0, PF_FLAG_VELOCITIES), 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    if (S.ct.on && t.first == 0 && threadIdx.x == 0) {
      // Substep start: the contact counter (resetCounterKernel, ContactBuffer.ts:62-65).
      __hip_atomic_store(S.ct.count, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    if (valid) {
      float4 p;
      const float w = S.w[parity][i];
      float4 vel;
      const float4 xsp = (w > 0.0f) ? make_float4(0.f, 0.f, 0.f, 0.f) : S.xs[xs_prev][i];
      predict(S, S.x[parity][i], S.v[parity][i], w, xsp, &vel, &p);
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
    if (over) pf_flag_raise(flag_word(S, parity, PF_FLAG_GRID));
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
template <bool PERSIST>
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
    S.tmp_id[slot] = S.id[parity][i];
  });
}

// Stable rank inside the bucket (ascending particle id), then the particle
// state of `parity` gathered into `parity ^ 1` in cell-sorted order,
// predicting velocities and positions (integrate.ts:25-34) on the way.
template <bool PERSIST>
__global__ __launch_bounds__(PF_BLOCK) void k_rank(const SceneDev* __restrict__ scenes,
                                                   const Tile* __restrict__ tiles, uint32_t ntiles,
                                                   int parity, int xs_prev, int xs_out) {
  pf_for_tiles<PERSIST>(ntiles, [&](uint32_t ti) {
    const Tile t = tiles[ti];
    const SceneDev& S = scenes[t.scene];
    if (t.first == 0 && threadIdx.x == 0) {
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
    const uint32_t* cmp = S.tmp_id;
    const uint32_t my = cmp[s];
    const uint32_t src = S.tmp_src[s];
    // The particle's state is gathered while the rank is counted.
    const float4 x = S.x[parity][src];
    const float4 v = S.v[parity][src];
    const float w = S.w[parity][src];
    float4 xsp = S.xs[xs_prev][src];
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
      pf_flag_raise(S.flags + parity * PF_NUM_FLAGS + PF_FLAG_LONG_BUCKET);
    }
    {
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
// Standalone rebuild (HashGrid.rebuild over a grid's own positions): an LSD
// radix sort of (bucket, index) pairs, 8 bits per pass.
//
// Points in arbitrary order (a uniform cloud, blobs) make the counting sort's
// histogram and placement random device-wide atomics, contended where buckets
// are dense. Here every pass ranks a 4096-point block in LDS and writes each
// digit's run to its place (the digit's global start plus the block's share,
// from per-block histograms scanned per digit), staged through LDS so that the
// writes of a run are contiguous. Each pass is stable, so the permutation is
// ascending index within a bucket, the same canonical order as the counting
// sort. Bucket offsets come from the sorted keys last: each bucket run's first
// position, then per 4096-bucket range a suffix minimum that gives empty
// buckets the next run's offset.
// ---------------------------------------------------------------------------
static_assert(PF_BLOCK == RS_BINS, "one thread per digit");

// Block-wide exclusive scan of one value per thread; `s_w` holds 4 words.
static __device__ __forceinline__ uint32_t rs_block_excl(uint32_t x, uint32_t* s_w, uint32_t* total) {
  const uint32_t lane = __lane_id();
  const uint32_t wave = threadIdx.x >> 6;
  uint32_t incl = x;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const uint32_t y = __shfl_up(incl, d);
    if (lane >= (uint32_t)d) incl += y;
  }
  if (lane == 63) s_w[wave] = incl;
  __syncthreads();
  uint32_t off = 0, tot = 0;
#pragma unroll
  for (uint32_t w = 0; w < PF_BLOCK / 64; ++w) {
    const uint32_t v = s_w[w];
    off += (w < wave) ? v : 0u;
    tot += v;
  }
  *total = tot;
  return off + incl - x;
}

// Keys of the grid's positions (cellIndex.ts:113-143) and the block
// histograms of their lowest digit. hist is digit-major: [d * nblk + block].
__global__ __launch_bounds__(PF_BLOCK) void k_rs_keys(const SceneDev* __restrict__ sc,
                                                      uint32_t* __restrict__ hist, uint32_t nblk) {
  __shared__ uint32_t s_h[RS_BINS];
  const SceneDev& S = *sc;
  const uint32_t blk = blockIdx.x;
  s_h[threadIdx.x] = 0u;
  __syncthreads();
  uint32_t over = 0;
#pragma unroll 4
  for (uint32_t u = 0; u < RS_ITEMS; ++u) {
    const uint32_t i = blk * RS_TILE + u * PF_BLOCK + threadIdx.x;
    if (i < S.n) {
      const float4 p = S.x[0][i];
      const uint32_t b = bucket_of(S, p.x, p.y, p.z, &over);
      S.key[i] = b;
      atomicAdd(&s_h[b & (RS_BINS - 1u)], 1u);
    }
  }
  __syncthreads();
  hist[threadIdx.x * nblk + blk] = s_h[threadIdx.x];
  if (over) pf_flag_raise(flag_word(S, 0, PF_FLAG_GRID));
}

// Block histograms of the digit at `shift` of keys in their current order.
__global__ __launch_bounds__(PF_BLOCK) void k_rs_hist(const uint32_t* __restrict__ keys, uint32_t n,
                                                      uint32_t shift, uint32_t* __restrict__ hist,
                                                      uint32_t nblk) {
  __shared__ uint32_t s_h[RS_BINS];
  const uint32_t blk = blockIdx.x;
  s_h[threadIdx.x] = 0u;
  uint32_t k[RS_ITEMS];
#pragma unroll
  for (uint32_t u = 0; u < RS_ITEMS; ++u) {
    const uint32_t i = blk * RS_TILE + u * PF_BLOCK + threadIdx.x;
    k[u] = i < n ? keys[i] : 0u;
  }
  __syncthreads();
#pragma unroll
  for (uint32_t u = 0; u < RS_ITEMS; ++u) {
    const uint32_t i = blk * RS_TILE + u * PF_BLOCK + threadIdx.x;
    if (i < n) atomicAdd(&s_h[(k[u] >> shift) & (RS_BINS - 1u)], 1u);
  }
  __syncthreads();
  hist[threadIdx.x * nblk + blk] = s_h[threadIdx.x];
}

// One workgroup per digit: its block counts become exclusive offsets within
// the digit, and the digit's total goes to hist[RS_BINS * nblk + d]. With
// `start` (the first pass) the workgroups also mark every bucket offset empty
// for k_rs_heads.
__global__ __launch_bounds__(PF_BLOCK) void k_rs_scan(uint32_t* __restrict__ hist, uint32_t nblk,
                                                      uint32_t* __restrict__ start, uint32_t H) {
  __shared__ uint32_t s_w[PF_BLOCK / 64];
  const uint32_t d = blockIdx.x;
  if (start)
    for (uint32_t i = d * PF_BLOCK + threadIdx.x; i < H; i += RS_BINS * PF_BLOCK) start[i] = 0xffffffffu;
  uint32_t* row = hist + (size_t)d * nblk;
  uint32_t carry = 0;
  for (uint32_t b0 = 0; b0 < nblk; b0 += PF_BLOCK) {
    const uint32_t b = b0 + threadIdx.x;
    const uint32_t c = b < nblk ? row[b] : 0u;
    uint32_t tot;
    const uint32_t ex = rs_block_excl(c, s_w, &tot);
    if (b < nblk) row[b] = carry + ex;
    carry += tot;
    __syncthreads();  // s_w is reused by the next chunk
  }
  if (threadIdx.x == 0) hist[(size_t)RS_BINS * nblk + d] = carry;
}

// One pass: rank the block's points by the digit at `shift` (stable: waves
// take consecutive quarters of the block, lanes consecutive points, items in
// order), stage them in LDS in block digit order, write each digit's run to
// its global place. FIRST: values are the points' indices. LAST with spos:
// also gathers the sorted positions.
template <bool FIRST, bool LAST>
__global__ __launch_bounds__(PF_BLOCK) void k_rs_scatter(
    const uint32_t* __restrict__ keys_in, const uint32_t* __restrict__ vals_in,
    uint32_t* __restrict__ keys_out, uint32_t* __restrict__ vals_out, uint32_t n, uint32_t shift,
    uint32_t dbits, const uint32_t* __restrict__ hist, uint32_t nblk, const float4* __restrict__ pos,
    float4* __restrict__ spos) {
  __shared__ uint32_t s_cnt[PF_BLOCK / 64][RS_BINS];
  __shared__ uint32_t s_bstart[RS_BINS];
  __shared__ uint32_t s_gbase[RS_BINS];
  __shared__ uint32_t s_w0[PF_BLOCK / 64], s_w1[PF_BLOCK / 64];
  __shared__ uint32_t s_k[RS_TILE], s_v[RS_TILE];
  const uint32_t t = threadIdx.x;
  const uint32_t lane = __lane_id();
  const uint32_t wave = t >> 6;
  const uint32_t blk = blockIdx.x;
  const unsigned long long lt = (1ull << lane) - 1ull;
#pragma unroll
  for (uint32_t w = 0; w < PF_BLOCK / 64; ++w) s_cnt[w][t] = 0u;
  const uint32_t digit_total = hist[(size_t)RS_BINS * nblk + t];
  const uint32_t block_off = hist[(size_t)t * nblk + blk];
  __syncthreads();

  uint32_t k[RS_ITEMS], v[RS_ITEMS], r[RS_ITEMS];
  const uint32_t wbase = blk * RS_TILE + wave * (RS_TILE / 4);
#pragma unroll
  for (uint32_t u = 0; u < RS_ITEMS; ++u) {
    const uint32_t e = wbase + u * 64u + lane;
    const bool valid = e < n;
    k[u] = valid ? keys_in[e] : 0u;
    v[u] = FIRST ? e : (valid ? vals_in[e] : 0u);
  }
#pragma unroll
  for (uint32_t u = 0; u < RS_ITEMS; ++u) {
    const uint32_t e = wbase + u * 64u + lane;
    const bool valid = e < n;
    const uint32_t d = (k[u] >> shift) & (RS_BINS - 1u);
    unsigned long long peers = __ballot(valid);
    for (uint32_t b = 0; b < dbits; ++b) {
      const bool bit = (d >> b) & 1u;
      const unsigned long long m = __ballot(bit);
      peers &= bit ? m : ~m;
    }
    uint32_t rank = 0;
    if (valid) {
      const uint32_t pre = s_cnt[wave][d];
      rank = pre + (uint32_t)__popcll(peers & lt);
      if ((peers & lt) == 0ull) s_cnt[wave][d] = pre + (uint32_t)__popcll(peers);
    }
    r[u] = rank;
  }
  __syncthreads();
  // Digit t: the waves' offsets within the block's run, the run's block-local
  // start, and its global start for this block.
  const uint32_t c0 = s_cnt[0][t], c1 = s_cnt[1][t], c2 = s_cnt[2][t], c3 = s_cnt[3][t];
  uint32_t btot, gtot;
  const uint32_t bstart = rs_block_excl(c0 + c1 + c2 + c3, s_w0, &btot);
  const uint32_t gstart = rs_block_excl(digit_total, s_w1, &gtot);
  s_cnt[0][t] = 0u;
  s_cnt[1][t] = c0;
  s_cnt[2][t] = c0 + c1;
  s_cnt[3][t] = c0 + c1 + c2;
  s_bstart[t] = bstart;
  s_gbase[t] = gstart + block_off - bstart;
  __syncthreads();
#pragma unroll
  for (uint32_t u = 0; u < RS_ITEMS; ++u) {
    const uint32_t e = wbase + u * 64u + lane;
    if (e < n) {
      const uint32_t d = (k[u] >> shift) & (RS_BINS - 1u);
      const uint32_t p = s_bstart[d] + s_cnt[wave][d] + r[u];
      s_k[p] = k[u];
      s_v[p] = v[u];
    }
  }
  __syncthreads();
  const uint32_t count = min(RS_TILE, n - blk * RS_TILE);
#pragma unroll 4
  for (uint32_t q = t; q < count; q += PF_BLOCK) {
    const uint32_t kk = s_k[q];
    const uint32_t vv = s_v[q];
    const uint32_t dest = s_gbase[(kk >> shift) & (RS_BINS - 1u)] + q;
    keys_out[dest] = kk;
    vals_out[dest] = vv;
    if (LAST && spos) spos[dest] = pos[vv];
  }
}

// Per sorted position q: the first of its bucket's run writes the bucket's
// offset (start[key] = q), and q writes the first position of the
// RS_RANGE-bucket ranges that begin after its predecessor's key and at or
// before its own (position n closes the ranges past the last key).
__global__ __launch_bounds__(PF_BLOCK) void k_rs_heads(const uint32_t* __restrict__ skey, uint32_t n,
                                                       uint32_t nranges, uint32_t* __restrict__ start,
                                                       uint32_t* __restrict__ rlo) {
  const uint32_t blk = blockIdx.x;
  const uint32_t base = blk * RS_TILE + threadIdx.x;
  uint32_t k[RS_ITEMS], before[RS_ITEMS];
  // every load in flight before any is used
#pragma unroll
  for (uint32_t u = 0; u < RS_ITEMS; ++u) {
    const uint32_t q = base + u * PF_BLOCK;
    k[u] = q < n ? skey[q] : 0u;
    before[u] = (q > 0 && q < n) ? skey[q - 1] : 0u;
  }
#pragma unroll
  for (uint32_t u = 0; u < RS_ITEMS; ++u) {
    const uint32_t q = base + u * PF_BLOCK;
    if (q < n) {
      if (q == 0 || before[u] != k[u]) start[k[u]] = q;
      const int pr = q == 0 ? -1 : (int)(before[u] / RS_RANGE);
      for (int r = pr + 1; r <= (int)(k[u] / RS_RANGE); ++r) rlo[r] = q;
    }
  }
  if (blk == gridDim.x - 1 && threadIdx.x == 0)
    for (uint32_t r = skey[n - 1] / RS_RANGE + 1; r <= nranges; ++r) rlo[r] = n;
}

// Bucket offsets of one range: an empty bucket's offset is the next non-empty
// bucket's, so each range takes a suffix minimum over its offsets, carrying in
// the next range's first position. The last range also writes start[H] = n.
__global__ __launch_bounds__(PF_BLOCK) void k_rs_ranges(const uint32_t* __restrict__ rlo, uint32_t n,
                                                        uint32_t H, uint32_t* __restrict__ start) {
  constexpr uint32_t PER = RS_RANGE / PF_BLOCK;
  __shared__ uint32_t s_w[PF_BLOCK / 64];
  const uint32_t r = blockIdx.x;
  const uint32_t lane = __lane_id();
  const uint32_t wave = threadIdx.x >> 6;
  const uint32_t carry = rlo[r + 1];
  const uint32_t base = r * RS_RANGE + threadIdx.x * PER;
  const bool whole = base + PER <= H;
  uint32_t v[PER];
  if (whole) {
    const uint4* src = reinterpret_cast<const uint4*>(start + base);
#pragma unroll
    for (uint32_t j = 0; j < PER; j += 4) {
      const uint4 q = src[j / 4];
      v[j] = q.x; v[j + 1] = q.y; v[j + 2] = q.z; v[j + 3] = q.w;
    }
  } else {
#pragma unroll
    for (uint32_t j = 0; j < PER; ++j) v[j] = base + j < H ? start[base + j] : 0xffffffffu;
  }
  uint32_t m = v[0];
#pragma unroll
  for (uint32_t j = 1; j < PER; ++j) m = min(m, v[j]);
  // inclusive suffix minimum over the wave's lanes, then over the waves
  uint32_t x = m;
#pragma unroll
  for (uint32_t d = 1; d < 64; d <<= 1) {
    const uint32_t y = __shfl_down(x, d);
    if (lane + d < 64u) x = min(x, y);
  }
  if (lane == 0) s_w[wave] = x;
  __syncthreads();
  uint32_t run = carry;
#pragma unroll
  for (uint32_t w = 0; w < PF_BLOCK / 64; ++w)
    if (w > wave) run = min(run, s_w[w]);
  const uint32_t next = __shfl_down(x, 1);
  if (lane < 63u) run = min(run, next);
#pragma unroll
  for (int j = (int)PER - 1; j >= 0; --j) {
    run = min(run, v[j]);
    v[j] = run;
  }
  if (whole) {
    uint4* dst = reinterpret_cast<uint4*>(start + base);
#pragma unroll
    for (uint32_t j = 0; j < PER; j += 4) dst[j / 4] = make_uint4(v[j], v[j + 1], v[j + 2], v[j + 3]);
  } else {
#pragma unroll
    for (uint32_t j = 0; j < PER; ++j)
      if (base + j < H) start[base + j] = v[j];
  }
  if (r == gridDim.x - 1 && threadIdx.x == 0) start[H] = n;
}

// ---------------------------------------------------------------------------
// Host launchers
// ---------------------------------------------------------------------------
void launch_grid_radix(hipStream_t st, const SceneDev* dev, const SceneDev& host, uint32_t* skey,
                       uint32_t* perm, float4* spos, uint32_t* scratch) {
  const uint32_t n = host.n, H = host.H;
  if (n == 0) return;
  const uint32_t nblk = (n + RS_TILE - 1) / RS_TILE;
  const uint32_t nranges = (H + RS_RANGE - 1) / RS_RANGE;
  uint32_t* hist = scratch;
  uint32_t* rlo = scratch + (size_t)RS_BINS * nblk + RS_BINS;
  const uint32_t bits = 31u - (uint32_t)__builtin_clz(H);  // H is a power of two
  const uint32_t passes = bits == 0 ? 1u : (bits + 7u) / 8u;
  hipLaunchKernelGGL(k_rs_keys, dim3(nblk), dim3(PF_BLOCK), 0, st, dev, hist, nblk);
  const uint32_t* kin = host.key;
  const uint32_t* vin = nullptr;
  for (uint32_t p = 0; p < passes; ++p) {
    const uint32_t shift = 8u * p;
    const uint32_t dbits = std::min(8u, bits > shift ? bits - shift : 0u);
    // the last pass lands in (skey, perm); earlier ones alternate with the tmp pair
    const bool to_final = ((passes - 1u - p) & 1u) == 0u;
    uint32_t* kout = to_final ? skey : host.tmp_key;
    uint32_t* vout = to_final ? perm : host.tmp_src;
    if (p > 0) hipLaunchKernelGGL(k_rs_hist, dim3(nblk), dim3(PF_BLOCK), 0, st, kin, n, shift, hist, nblk);
    hipLaunchKernelGGL(k_rs_scan, dim3(RS_BINS), dim3(PF_BLOCK), 0, st, hist, nblk, p == 0 ? host.start : nullptr,
                       H);
    const bool first = p == 0, last = p + 1 == passes;
    auto kern = first ? (last ? k_rs_scatter<true, true> : k_rs_scatter<true, false>)
                      : (last ? k_rs_scatter<false, true> : k_rs_scatter<false, false>);
    hipLaunchKernelGGL(kern, dim3(nblk), dim3(PF_BLOCK), 0, st, kin, vin, kout, vout, n, shift, dbits,
                       (const uint32_t*)hist, nblk, (const float4*)host.x[0], last ? spos : nullptr);
    kin = kout;
    vin = vout;
  }
  hipLaunchKernelGGL(k_rs_heads, dim3(nblk), dim3(PF_BLOCK), 0, st, (const uint32_t*)skey, n, nranges,
                     host.start, rlo);
  hipLaunchKernelGGL(k_rs_ranges, dim3(nranges), dim3(PF_BLOCK), 0, st, (const uint32_t*)rlo, n, H, host.start);
}

void launch_hash(hipStream_t st, const SceneDev* scenes, const Tile* tiles, uint32_t ntiles, int parity,
                 int xs_prev, int reset_step_flags) {
  launch_tiles(k_hash<false>, k_hash<true>, ntiles, 0, st, scenes, tiles, ntiles, parity, xs_prev,
               reset_step_flags);
}

void launch_scan(hipStream_t st, const SceneDev* scenes, const uint2* stiles, uint32_t nstiles,
                 void* ctl) {
  if (nstiles == 0) return;
  hipLaunchKernelGGL(k_scan, dim3(nstiles), dim3(PF_BLOCK), 0, st, scenes, stiles, nstiles,
                     reinterpret_cast<ScanCtlDev*>(ctl));
}

void launch_place(hipStream_t st, const SceneDev* scenes, const Tile* tiles, uint32_t ntiles, int parity) {
  launch_tiles(k_place<false>, k_place<true>, ntiles, 0, st, scenes, tiles, ntiles, parity);
}

void launch_rank(hipStream_t st, const SceneDev* scenes, const Tile* tiles, uint32_t ntiles, int parity,
                 int xs_prev, int xs_out) {
  launch_tiles(k_rank<false>, k_rank<true>, ntiles, 0, st, scenes, tiles, ntiles, parity, xs_prev, xs_out);
}

}  // namespace pf
