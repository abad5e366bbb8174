// Hash-grid counting sort that allocates nothing: pf_grid_sort_noalloc.
//
// The same four sort kernels as pf_grid_rebuild over a standalone grid
// (k_hash, k_scan, k_place, k_rank in grid.hip), for callers that cannot
// allocate device memory (code under HIP graph capture, tight serving loops).
// Inputs and outputs are the caller's device buffers; everything else the
// kernels need (the grid descriptor, the work tiles, arrival ranks, the
// arrival-order scatter, the bucket offsets, the scan's status words and
// ticket, the overflow flags) lives in one file-scope device array,
// `scratch`, of fixed size. Each GPU holds its own copy of that array until
// the process ends; its address on a device is taken once, with that device
// current, and every launch is given that address. Calls that overlap on one
// GPU share the array and race; successive calls on one stream do not.
#include <algorithm>
#include <cmath>
#include <mutex>

#include "runtime.hpp"

namespace pf {
namespace {

constexpr uint32_t kMaxN = PF_NOALLOC_MAX_PARTICLES;
constexpr uint32_t kMaxH = PF_NOALLOC_MAX_BUCKETS;
constexpr size_t kAlign = 256;
constexpr size_t align_up(size_t v) { return (v + kAlign - 1) / kAlign * kAlign; }

constexpr uint32_t kMaxTiles = kMaxN / PF_BLOCK;
constexpr uint32_t kMaxScanTiles = kMaxH / PF_SCAN_TILE;

// Byte offsets inside the array.
constexpr size_t kDesc = 0;
constexpr size_t kCtl = align_up(kDesc + sizeof(SceneDev));  // scan ticket (16 B), scan_ctl (8 B), flags
constexpr size_t kCtlBytes = 64 + 4 * 2 * PF_NUM_FLAGS;
constexpr size_t kStatus = align_up(kCtl + kCtlBytes);
constexpr size_t kTiles = align_up(kStatus + 8 * (size_t)kMaxScanTiles);
constexpr size_t kSTiles = align_up(kTiles + sizeof(Tile) * (size_t)kMaxTiles);
constexpr size_t kRank = align_up(kSTiles + sizeof(uint2) * (size_t)kMaxScanTiles);
constexpr size_t kSrc = align_up(kRank + 4 * (size_t)kMaxN);
constexpr size_t kStart = align_up(kSrc + 4 * (size_t)kMaxN);
constexpr size_t kTotal = align_up(kStart + 4 * ((size_t)kMaxH + 1));
static_assert(kTotal == PF_NOALLOC_SCRATCH_BYTES, "scratch size and its documented value differ");

}  // namespace

static __device__ __align__(256) unsigned char scratch[kTotal];

namespace {

// Writes the grid descriptor (block 0) and the work tiles, and clears the
// histogram. One launch ahead of the sort kernels on the same stream.
__global__ __launch_bounds__(PF_BLOCK) void k_noalloc_setup(unsigned char* base, const float4* positions,
                                                            uint32_t n, uint32_t H, float cell, float ox,
                                                            float oy, float oz, uint32_t* key,
                                                            uint32_t* counts, uint32_t* sorted_keys) {
  const uint32_t ntiles = (n + PF_BLOCK - 1) / PF_BLOCK;
  const uint32_t nstiles = (H + PF_SCAN_TILE - 1) / PF_SCAN_TILE;
  if (blockIdx.x == 0) {
    uint32_t* words = reinterpret_cast<uint32_t*>(base + kDesc);
    for (uint32_t i = threadIdx.x; i < sizeof(SceneDev) / 4; i += PF_BLOCK) words[i] = 0u;
    uint32_t* flags = reinterpret_cast<uint32_t*>(base + kCtl + 64);
    for (uint32_t i = threadIdx.x; i < 2 * PF_NUM_FLAGS; i += PF_BLOCK) flags[i] = 0u;
    __syncthreads();
    if (threadIdx.x == 0) {
      SceneDev* S = reinterpret_cast<SceneDev*>(base + kDesc);
      S->n = n;
      S->H = H;
      S->hash_mask = H - 1u;
      S->nscan_tiles = nstiles;
      S->cell = cell;
      S->ox = ox;
      S->oy = oy;
      S->oz = oz;
      S->x[0] = const_cast<float4*>(positions);
      S->key = key;
      S->rank = reinterpret_cast<uint32_t*>(base + kRank);
      S->counts = counts;
      S->start = reinterpret_cast<uint32_t*>(base + kStart);
      S->scan_status = reinterpret_cast<unsigned long long*>(base + kStatus);
      S->scan_ctl = reinterpret_cast<uint32_t*>(base + kCtl + 16);
      S->tmp_src = reinterpret_cast<uint32_t*>(base + kSrc);
      S->tmp_key = sorted_keys;  // k_rank reads entry s and writes the sorted key s in place
      S->flags = flags;
    }
  }
  Tile* tiles = reinterpret_cast<Tile*>(base + kTiles);
  uint2* stiles = reinterpret_cast<uint2*>(base + kSTiles);
  const uint32_t stride = gridDim.x * PF_BLOCK;
  for (uint32_t i = blockIdx.x * PF_BLOCK + threadIdx.x; i < H; i += stride) {
    counts[i] = 0u;
    if (i < ntiles) tiles[i] = Tile{0u, i * PF_BLOCK};
    if (i < nstiles) stiles[i] = make_uint2(0u, i);
  }
  for (uint32_t i = H + blockIdx.x * PF_BLOCK + threadIdx.x; i < ntiles; i += stride)
    tiles[i] = Tile{0u, i * PF_BLOCK};
}

// HashGridSnapshot ranges from the bucket offsets (HashGrid.ts:55-64,
// padded to Hpad), and the overflow flag (HashGrid.ts:256-260).
__global__ __launch_bounds__(PF_BLOCK) void k_noalloc_ranges(const unsigned char* base, uint32_t Hpad,
                                                             uint32_t* counts, uint32_t* cell_start,
                                                             uint32_t* cell_end, uint32_t* overflow) {
  const SceneDev* S = reinterpret_cast<const SceneDev*>(base + kDesc);
  const uint32_t H = S->H;
  const uint32_t* start = S->start;
  const uint32_t total = start[H];
  for (uint32_t b = blockIdx.x * PF_BLOCK + threadIdx.x; b < Hpad; b += gridDim.x * PF_BLOCK) {
    const uint32_t s0 = b < H ? start[b] : total;
    const uint32_t s1 = b < H ? start[b + 1] : total;
    counts[b] = s1 - s0;
    cell_start[b] = s0;
    cell_end[b] = s1;
  }
  if (overflow && blockIdx.x == 0 && threadIdx.x == 0) overflow[0] = S->flags[PF_FLAG_GRID] != 0u ? 1u : 0u;
}

// The array's address on each device, taken with that device current.
std::mutex g_scratch_mu;
unsigned char* g_scratch_ptr[64] = {};

unsigned char* scratch_on(int device) {
  if (device < 0 || device >= 64) fail(PF_EDEVICE, "HashGrid.sortNoAlloc: device %d out of range", device);
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      fail(PF_EDEVICE, "HashGrid.sortNoAlloc: no address for the scratch array on device %d: %s", device,
           hipGetErrorString(e));
    g_scratch_ptr[device] = static_cast<unsigned char*>(p);
  }
  return g_scratch_ptr[device];
}

uint32_t next_pow2_u32(uint64_t v) {
  uint64_t p = 1;
  while (p < v) p <<= 1;
  return (uint32_t)p;
}

}  // namespace
}  // namespace pf

#define PF_API_BEGIN try {
#define PF_API_END                                    \
  }                                                   \
  catch (const pf::PfError& e) { return e.code; }     \
  catch (const pf::HipError& e) { return e.code; }    \
  catch (const std::exception& e) {                   \
    pf::set_error("%s", e.what());                    \
    return PF_EINVAL;                                 \
  }                                                   \
  return PF_OK;

extern "C" {

int pf_grid_sort_noalloc(pf_context* ctx, const float* positions4, uint32_t n, double cell_size,
                         uint32_t hash_table_size, const double origin[3], uint32_t* cell_index,
                         uint32_t* counts, uint32_t* cell_start, uint32_t* cell_end,
                         uint32_t* sorted_indices, uint32_t* sorted_keys, uint32_t* overflow) {
  using namespace pf;
  PF_API_BEGIN
  if (!ctx) fail(PF_EINVAL, "HashGrid.sortNoAlloc: null handle");
  if (!positions4 || !cell_index || !counts || !cell_start || !cell_end || !sorted_indices || !sorted_keys)
    fail(PF_EINVAL, "HashGrid.sortNoAlloc: positions and every output but overflow are required");
  if (reinterpret_cast<uintptr_t>(positions4) % 16u != 0)
    fail(PF_EINVAL, "HashGrid.sortNoAlloc: positions must be 16-byte aligned vec4 data");
  if (n == 0 || n > PF_NOALLOC_MAX_PARTICLES)
    fail(PF_EINVAL, "HashGrid.sortNoAlloc: %u positions; this entry point sorts 1 to %u (fixed scratch of %u bytes)",
         n, (uint32_t)PF_NOALLOC_MAX_PARTICLES, (uint32_t)PF_NOALLOC_SCRATCH_BYTES);
  if (!(cell_size > 0) || !std::isfinite(cell_size))
    fail(PF_EINVAL, "HashGrid: cellSize must be a positive finite number, got %g", cell_size);
  uint32_t H = hash_table_size;
  if (H == 0) H = std::min<uint32_t>(next_pow2_u32(2ull * n), 1u << 20);
  if ((H & (H - 1)) != 0) fail(PF_EINVAL, "HashGrid: hashTableSize must be a positive power of two, got %u", H);
  if (H > PF_NOALLOC_MAX_BUCKETS)
    fail(PF_EINVAL, "HashGrid.sortNoAlloc: hashTableSize %u exceeds this entry point's %u (fixed scratch)", H,
         (uint32_t)PF_NOALLOC_MAX_BUCKETS);
  const uint32_t Hpad = (uint32_t)(((uint64_t)H + 1023) / 1024 * 1024);
  double o[3] = {0.0, 0.0, 0.0};
  if (origin) for (int a = 0; a < 3; ++a) o[a] = origin[a];

  int prev = 0;
  PF_HIP(hipGetDevice(&prev));
  if (prev != ctx->device) PF_HIP(hipSetDevice(ctx->device));
  struct Restore {
    int dev, prev;
    ~Restore() {
      if (dev != prev) (void)hipSetDevice(prev);
    }
  } restore{ctx->device, prev};
  unsigned char* base = scratch_on(ctx->device);
  hipStream_t st = ctx->stream;
  const SceneDev* S = reinterpret_cast<const SceneDev*>(base + kDesc);
  Tile* tiles = reinterpret_cast<Tile*>(base + kTiles);
  uint2* stiles = reinterpret_cast<uint2*>(base + kSTiles);
  const uint32_t ntiles = (n + PF_BLOCK - 1) / PF_BLOCK;
  const uint32_t nstiles = (H + PF_SCAN_TILE - 1) / PF_SCAN_TILE;
  const uint32_t fill_blocks = std::min<uint32_t>((std::max(H, ntiles) + PF_BLOCK - 1) / PF_BLOCK, 2048u);
  hipLaunchKernelGGL(k_noalloc_setup, dim3(fill_blocks), dim3(PF_BLOCK), 0, st, base,
                     reinterpret_cast<const float4*>(positions4), n, H, (float)cell_size, (float)o[0],
                     (float)o[1], (float)o[2], cell_index, counts, sorted_keys);
  launch_hash(st, S, tiles, ntiles, 0, 0, 0, 0);
  launch_scan(st, S, stiles, nstiles, base + kCtl);
  launch_place(st, S, tiles, ntiles, 0, 0);
  launch_rank(st, S, tiles, ntiles, 0, 0, 0, 0, sorted_indices, sorted_keys, nullptr);
  hipLaunchKernelGGL(k_noalloc_ranges, dim3(std::min<uint32_t>((Hpad + PF_BLOCK - 1) / PF_BLOCK, 2048u)),
                     dim3(PF_BLOCK), 0, st, base, Hpad, counts, cell_start, cell_end, overflow);
  PF_HIP(hipGetLastError());
  PF_API_END
}

}  // extern "C"
