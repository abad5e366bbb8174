// Hash-grid sort that allocates nothing: pf_grid_sort_noalloc.
//
// The same radix-sort kernels as pf_grid_rebuild over a standalone grid
// (launch_grid_radix in grid.hip), for callers that cannot allocate device
// memory (code under HIP graph capture, tight serving loops). Inputs and
// outputs are the caller's device buffers; everything else the kernels need
// (the grid descriptor, the overflow flags, the passes' other key and index
// buffers, the bucket offsets, the digit histograms and range bounds) lives
// in one file-scope device array, `scratch`, of fixed size. Each GPU holds
// its own copy of that array until the process ends; its address on a device
// is taken once, with that device current, and every launch is given that
// address. Calls that overlap on one GPU share the array and race; successive
// calls on one stream do not.
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

// Byte offsets inside the array.
constexpr size_t kDesc = 0;
constexpr size_t kFlags = align_up(kDesc + sizeof(SceneDev));
constexpr size_t kTmpKey = align_up(kFlags + 4 * 2 * PF_NUM_FLAGS);
constexpr size_t kTmpSrc = align_up(kTmpKey + 4 * (size_t)kMaxN);
constexpr size_t kStart = align_up(kTmpSrc + 4 * (size_t)kMaxN);
constexpr size_t kRadix = align_up(kStart + 4 * ((size_t)kMaxH + 1));
constexpr size_t kTotal = align_up(kRadix + 4 * grid_radix_scratch_words(kMaxN, kMaxH));
static_assert(kTotal == PF_NOALLOC_SCRATCH_BYTES, "scratch size and its documented value differ");
// the descriptor travels by value as a kernel argument (at most 4 KiB of them)
static_assert(sizeof(SceneDev) + sizeof(void*) <= 4096, "descriptor too large for a kernel argument");

}  // namespace

static __device__ __align__(256) unsigned char scratch[kTotal];

namespace {

// The descriptor's device copy, with every pointer into the caller's buffers
// or the array, and cleared overflow flags. One launch ahead of the sort.
__global__ __launch_bounds__(PF_BLOCK) void k_noalloc_setup(unsigned char* base, SceneDev desc) {
  uint32_t* flags = reinterpret_cast<uint32_t*>(base + kFlags);
  for (uint32_t i = threadIdx.x; i < 2 * PF_NUM_FLAGS; i += PF_BLOCK) flags[i] = 0u;
  static_assert(sizeof(SceneDev) % 4 == 0, "descriptor copied as words");
  const uint32_t* src = reinterpret_cast<const uint32_t*>(&desc);
  uint32_t* dst = reinterpret_cast<uint32_t*>(base + kDesc);
  for (uint32_t i = threadIdx.x; i < sizeof(SceneDev) / 4; i += PF_BLOCK) dst[i] = src[i];
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
  SceneDev desc{};
  desc.n = n;
  desc.H = H;
  desc.hash_mask = H - 1u;
  desc.cell = (float)cell_size;
  desc.ox = (float)o[0];
  desc.oy = (float)o[1];
  desc.oz = (float)o[2];
  desc.x[0] = const_cast<float4*>(reinterpret_cast<const float4*>(positions4));
  desc.key = cell_index;
  desc.start = reinterpret_cast<uint32_t*>(base + kStart);
  desc.tmp_key = reinterpret_cast<uint32_t*>(base + kTmpKey);
  desc.tmp_src = reinterpret_cast<uint32_t*>(base + kTmpSrc);
  desc.flags = reinterpret_cast<uint32_t*>(base + kFlags);
  const SceneDev* S = reinterpret_cast<const SceneDev*>(base + kDesc);
  hipLaunchKernelGGL(k_noalloc_setup, dim3(1), dim3(PF_BLOCK), 0, st, base, desc);
  launch_grid_radix(st, S, desc, sorted_keys, sorted_indices, nullptr, reinterpret_cast<uint32_t*>(base + kRadix));
  hipLaunchKernelGGL(k_noalloc_ranges, dim3(std::min<uint32_t>((Hpad + PF_BLOCK - 1) / PF_BLOCK, 2048u)),
                     dim3(PF_BLOCK), 0, st, base, Hpad, counts, cell_start, cell_end, overflow);
  PF_HIP(hipGetLastError());
  PF_API_END
}

}  // extern "C"
