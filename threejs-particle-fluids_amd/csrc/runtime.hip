// Context, ParticleSystem and HashGrid behind the C ABI, plus the small
// gather/scatter kernels that move particle data between original order
// (the ABI) and cell-sorted order (the device).
#include <sys/mman.h>

#include <algorithm>
#include <cmath>
#include <condition_variable>
#include <cstdlib>
#include <mutex>
#include <stdexcept>
#include <thread>
#include <unordered_set>

#include "runtime.hpp"

namespace pf {

static thread_local std::string g_error;


void set_error(const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_error = buf;
}

void fail(int code, const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_error = buf;
  throw PfError{code};
}

// ---------------------------------------------------------------------------
// Device allocation: hipMalloc, or a scene arena inside an ArenaScope.
// ---------------------------------------------------------------------------
static std::mutex g_alloc_mu;
static std::unordered_set<void*> g_malloced;  // buffers dfree may hand to hipFree
static thread_local Arena* g_arena = nullptr;

static constexpr size_t kArenaMinChunk = 4u << 20;
static constexpr size_t kArenaBytesPerParticle = 1024;
static constexpr size_t kArenaPage = 2u << 20;
static constexpr size_t kArenaAlign = 256;
// Mapped tail no buffer uses: reads a little past a buffer's end stay inside
// mapped memory, as they do with hipMalloc's rounded allocations.
static constexpr size_t kArenaTail = 64u << 10;

static hipMemAllocationProp arena_prop(int device) {
  hipMemAllocationProp prop{};
  prop.type = hipMemAllocationTypePinned;
  prop.location.type = hipMemLocationTypeDevice;
  prop.location.id = device;
  return prop;
}

// Backs a segment's whole address range with one new physical allocation.
static void arena_map(Arena::Segment& g, int device) {
  const hipMemAllocationProp prop = arena_prop(device);
  hipError_t e = hipMemCreate(&g.handle, g.size, &prop, 0);
  if (e != hipSuccess) fail(PF_ENOMEM, "device allocation of %zu bytes failed: %s", g.size, hipGetErrorString(e));
  e = hipMemMap(g.base, g.size, 0, g.handle, 0);
  if (e != hipSuccess) {
    (void)hipMemRelease(g.handle);
    fail(PF_EDEVICE, "HIP: mapping %zu bytes at %p failed: %s", g.size, (void*)g.base, hipGetErrorString(e));
  }
  hipMemAccessDesc acc{};
  acc.location = prop.location;
  acc.flags = hipMemAccessFlagsProtReadWrite;
  e = hipMemSetAccess(g.base, g.size, &acc, 1);
  if (e != hipSuccess) {
    (void)hipMemUnmap(g.base, g.size);
    (void)hipMemRelease(g.handle);
    fail(PF_EDEVICE, "HIP: access to %zu bytes at %p failed: %s", g.size, (void*)g.base, hipGetErrorString(e));
  }
}

static void arena_unmap(Arena::Segment& g) {
  (void)hipMemUnmap(g.base, g.size);
  (void)hipMemRelease(g.handle);
}

Arena::Arena(int dev, size_t hint_bytes) : device(dev), hint(hint_bytes) {
  int vmm = 0;
  PF_HIP(hipDeviceGetAttribute(&vmm, hipDeviceAttributeVirtualMemoryManagementSupported, dev));
  if (!vmm) fail(PF_EDEVICE, "ParticleSystem: device %d has no virtual memory management; evictable scenes need it", dev);
  const hipMemAllocationProp prop = arena_prop(dev);
  PF_HIP(hipMemGetAllocationGranularity(&gran, &prop, hipMemAllocationGranularityRecommended));
  if (gran == 0) PF_HIP(hipMemGetAllocationGranularity(&gran, &prop, hipMemAllocationGranularityMinimum));
  if (gran == 0) gran = 4096;
}

Arena::~Arena() {
  for (auto& g : segs) {
    if (resident) arena_unmap(g);
    std::free(g.host);
    (void)hipMemAddressFree(g.base, g.size);
  }
}

// Evicted chunks live in pageable host memory; they move through a fixed
// ring of two pinned staging buffers per device (allocated once), so the
// pinned memory an eviction needs does not grow with the scenes evicted.
// One piece's DMA overlaps the host copy of the piece before it, and each
// host copy is split over a few worker threads (a single thread's memcpy,
// page faults included, runs well below the DMA's rate).
namespace {

// Host memory for an evicted segment: 2 MiB-aligned and marked for huge
// pages where the kernel allows it, so writing it takes few page faults.
char* host_alloc(size_t n) {
  constexpr size_t kHuge = 2u << 20;
  const size_t m = std::max<size_t>((n + kHuge - 1) / kHuge * kHuge, kHuge);
  void* p = std::aligned_alloc(kHuge, m);
  if (p) (void)madvise(p, m, MADV_HUGEPAGE);
  return static_cast<char*>(p);
}

class CopyPool {
 public:
  static constexpr int kWorkers = 4;
  CopyPool() {
    for (int i = 0; i < kWorkers; ++i) threads_.emplace_back([this, i] { run(i); });
  }
  ~CopyPool() {
    {
      std::lock_guard<std::mutex> lk(mu_);
      stop_ = true;
    }
    cv_.notify_all();
    for (auto& t : threads_) t.join();
  }
  // memcpy(dst, src, n), the calling thread taking one share.
  void copy(char* dst, const char* src, size_t n) {
    if (n < (4u << 20)) {
      std::memcpy(dst, src, n);
      return;
    }
    const size_t share = (n / (kWorkers + 1) + 4095) / 4096 * 4096;
    {
      std::lock_guard<std::mutex> lk(mu_);
      dst_ = dst;
      src_ = src;
      n_ = n;
      share_ = share;
      pending_ = kWorkers;
      ++gen_;
    }
    cv_.notify_all();
    const size_t off = share * kWorkers;
    if (off < n) std::memcpy(dst + off, src + off, n - off);
    std::unique_lock<std::mutex> lk(mu_);
    done_.wait(lk, [this] { return pending_ == 0; });
  }

 private:
  void run(int i) {
    uint64_t seen = 0;
    for (;;) {
      char* dst;
      const char* src;
      size_t n, share;
      {
        std::unique_lock<std::mutex> lk(mu_);
        cv_.wait(lk, [&] { return stop_ || gen_ != seen; });
        if (stop_) return;
        seen = gen_;
        dst = dst_;
        src = src_;
        n = n_;
        share = share_;
      }
      const size_t off = share * (size_t)i;
      if (off < n) std::memcpy(dst + off, src + off, std::min(share, n - off));
      {
        std::lock_guard<std::mutex> lk(mu_);
        if (--pending_ == 0) done_.notify_one();
      }
    }
  }
  std::mutex mu_;
  std::condition_variable cv_, done_;
  std::vector<std::thread> threads_;
  char* dst_ = nullptr;
  const char* src_ = nullptr;
  size_t n_ = 0, share_ = 0;
  int pending_ = 0;
  uint64_t gen_ = 0;
  bool stop_ = false;
};

struct StagingRing {
  static constexpr size_t kBytes = 32u << 20;
  std::mutex mu;
  char* buf[2] = {nullptr, nullptr};
  hipEvent_t ev[2] = {nullptr, nullptr};
  hipStream_t stream = nullptr;
  std::unique_ptr<CopyPool> pool;

  void init() {
    if (stream) return;
    for (int k = 0; k < 2; ++k) {
      PF_HIP(hipHostMalloc(reinterpret_cast<void**>(&buf[k]), kBytes, hipHostMallocDefault));
      PF_HIP(hipEventCreateWithFlags(&ev[k], hipEventDisableTiming));
    }
    PF_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    pool.reset(new CopyPool());
  }
  // Device -> pageable host.
  void download(char* dst, const char* src, size_t n) {
    size_t off = 0, pend_off = 0, pend_n = 0;
    int slot = 0, pend = -1;
    while (off < n || pend >= 0) {
      int issued = -1;
      size_t m = 0;
      if (off < n) {
        m = std::min(kBytes, n - off);
        PF_HIP(hipMemcpyAsync(buf[slot], src + off, m, hipMemcpyDeviceToHost, stream));
        PF_HIP(hipEventRecord(ev[slot], stream));
        issued = slot;
      }
      if (pend >= 0) {
        PF_HIP(hipEventSynchronize(ev[pend]));
        pool->copy(dst + pend_off, buf[pend], pend_n);
      }
      pend = issued;
      pend_off = off;
      pend_n = m;
      off += m;
      slot ^= 1;
    }
  }
  // Pageable host -> device; returns once the device holds the data.
  void upload(char* dst, const char* src, size_t n) {
    int slot = 0;
    for (size_t off = 0; off < n; off += kBytes) {
      const size_t m = std::min(kBytes, n - off);
      PF_HIP(hipEventSynchronize(ev[slot]));  // the copy that last read this buffer is done
      pool->copy(buf[slot], src + off, m);
      PF_HIP(hipMemcpyAsync(dst + off, buf[slot], m, hipMemcpyHostToDevice, stream));
      PF_HIP(hipEventRecord(ev[slot], stream));
      slot ^= 1;
    }
    PF_HIP(hipStreamSynchronize(stream));
  }
};

std::mutex g_ring_mu;
std::map<int, std::unique_ptr<StagingRing>> g_rings;

StagingRing& ring_for(int device) {
  std::lock_guard<std::mutex> lk(g_ring_mu);
  auto& r = g_rings[device];
  if (!r) r.reset(new StagingRing());
  return *r;
}
}  // namespace

// Buffers are bump-allocated in segments. The first is sized by the
// scene's expected bytes, each later one at least a quarter of what the
// arena holds already, so a scene holds few physical allocations (the device
// caps their number at about 32k per process, far below its memory: 32k
// chunks of 4 MiB stop at 128 GiB).
void* Arena::alloc(size_t bytes) {
  if (!resident) fail(PF_ESTATE, "ParticleSystem: the scene is evicted to host memory");
  bytes = (bytes + kArenaAlign - 1) / kArenaAlign * kArenaAlign;
  if (!segs.empty() && segs.back().size - segs.back().used >= bytes + kArenaTail) {
    Segment& g = segs.back();
    void* p = g.base + g.used;
    g.used += bytes;
    return p;
  }
  Segment g;
  const size_t want = std::max({bytes + kArenaTail, kArenaMinChunk, segs.empty() ? hint : this->bytes() / 4});
  // whole 2 MiB pages: the device backs large allocations with them
  const size_t page = std::max<size_t>(gran, kArenaPage);
  g.size = (want + page - 1) / page * page;
  void* va = nullptr;
  const hipError_t e = hipMemAddressReserve(&va, g.size, gran, nullptr, 0);
  if (e != hipSuccess) fail(PF_ENOMEM, "device address reservation of %zu bytes failed: %s", g.size, hipGetErrorString(e));
  g.base = static_cast<char*>(va);
  try {
    arena_map(g, device);
  } catch (...) {
    (void)hipMemAddressFree(g.base, g.size);
    throw;
  }
  g.used = bytes;
  segs.push_back(g);
  return g.base;
}

size_t Arena::bytes() const {
  size_t b = 0;
  for (const auto& g : segs) b += g.size;
  return b;
}

size_t Arena::used_bytes() const {
  size_t b = 0;
  for (const auto& g : segs) b += g.used;
  return b;
}

// Copies each segment's used bytes to pageable host memory, then releases
// the physical memory. Any failure (no host memory, a copy error) frees the
// partial copies and leaves the scene resident and intact.
void Arena::evict() {
  if (!resident) return;
  std::vector<char*> copies(segs.size(), nullptr);
  try {
    for (size_t k = 0; k < segs.size(); ++k) {
      copies[k] = host_alloc(segs[k].used);
      if (!copies[k]) fail(PF_ENOMEM, "Server: no host memory for an evicted scene's %zu bytes", segs[k].used);
    }
    StagingRing& R = ring_for(device);
    std::lock_guard<std::mutex> lk(R.mu);
    R.init();
    for (size_t k = 0; k < segs.size(); ++k) R.download(copies[k], segs[k].base, segs[k].used);
  } catch (...) {
    for (char* p : copies) std::free(p);
    throw;
  }
  for (size_t k = 0; k < segs.size(); ++k) {
    segs[k].host = copies[k];
    arena_unmap(segs[k]);
  }
  resident = false;
}

// Backs each segment with a fresh physical allocation at the same addresses
// and copies the host copies back. On failure every new mapping is released
// and the scene stays evicted, its host copies intact.
void Arena::restore() {
  if (resident) return;
  size_t mapped = 0;
  try {
    for (; mapped < segs.size(); ++mapped) arena_map(segs[mapped], device);
    StagingRing& R = ring_for(device);
    std::lock_guard<std::mutex> lk(R.mu);
    R.init();
    for (auto& g : segs) R.upload(g.base, static_cast<const char*>(g.host), g.used);
  } catch (...) {
    for (size_t k = 0; k < mapped; ++k) arena_unmap(segs[k]);
    throw;
  }
  for (auto& g : segs) {
    std::free(g.host);
    g.host = nullptr;
  }
  resident = true;
}

ArenaScope::ArenaScope(Arena* arena) : prev(g_arena) { g_arena = arena; }
ArenaScope::~ArenaScope() { g_arena = prev; }
Arena* current_arena() { return g_arena; }

void* dalloc_bytes(size_t bytes, bool temporary) {
  void* p = nullptr;
  if (g_arena && !temporary) {
    p = g_arena->alloc(bytes);
  } else {
    hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess) fail(PF_ENOMEM, "device allocation of %zu bytes failed: %s", bytes, hipGetErrorString(e));
    std::lock_guard<std::mutex> lk(g_alloc_mu);
    g_malloced.insert(p);
  }
  PF_HIP(hipMemset(p, 0, bytes));
  // Library streams are non-blocking: finish the clear before any of them runs.
  PF_HIP(hipDeviceSynchronize());
  return p;
}

void dfree(void* p) {
  if (!p) return;
  {
    std::lock_guard<std::mutex> lk(g_alloc_mu);
    if (g_malloced.erase(p) == 0) return;  // arena buffer
  }
  (void)hipFree(p);
}

hipEvent_t Context::get_event() {
  if (!event_pool.empty()) {
    hipEvent_t e = event_pool.back();
    event_pool.pop_back();
    return e;
  }
  hipEvent_t e;
  PF_HIP(hipEventCreate(&e));
  return e;
}

void Context::collect() {
  if (pending.empty()) return;
  PF_HIP(hipStreamSynchronize(stream));
  for (auto& p : pending) {
    float ms = 0.f;
    PF_HIP(hipEventElapsedTime(&ms, p.second.first, p.second.second));
    KernelStat& ks = stats[p.first];
    ks.launches += 1;
    ks.total_ms += ms;
    event_pool.push_back(p.second.first);
    event_pool.push_back(p.second.second);
  }
  pending.clear();
}

LaunchScope::LaunchScope(Context* c, const char* n) : ctx(c), name(n) {
  if (ctx->profiling && !ctx->capturing) {
    e0 = ctx->get_event();
    (void)hipEventRecord(e0, ctx->stream);
  }
}
LaunchScope::~LaunchScope() {
  if (e0) {
    e1 = ctx->get_event();
    (void)hipEventRecord(e1, ctx->stream);
    ctx->pending.push_back({name, {e0, e1}});
    if (ctx->pending.size() > 4096) {
      try { ctx->collect(); } catch (...) {}
    }
  }
}

SceneDev Particles::desc(int v) const {
  SceneDev d = host;
  if (v & 2) {
    std::swap(d.x[0], d.x[1]); std::swap(d.v[0], d.v[1]); std::swap(d.w[0], d.w[1]);
    std::swap(d.id[0], d.id[1]); std::swap(d.rho[0], d.rho[1]); std::swap(d.bvol[0], d.bvol[1]);
    std::swap(d.grp[0], d.grp[1]);
  }
  if (v & 1) std::swap(d.xs[0], d.xs[1]);
  return d;
}

void Particles::upload_desc() {
  if (!dirty) return;
  PF_HIP(hipStreamSynchronize(ctx->stream));
  SceneDev d[4] = {desc(0), desc(1), desc(2), desc(3)};
  PF_HIP(hipMemcpy(dev, d, sizeof d, hipMemcpyHostToDevice));
  dirty = false;
}

// ---------------------------------------------------------------------------
// Utility kernels: original order <-> cell-sorted order
// ---------------------------------------------------------------------------
enum GatherWhat {
  G_POS = 0, G_PRED = 1, G_VEL = 2, G_INVMASS = 3, G_DENSITY = 4, G_KEY = 5, G_IDS = 6, G_BVOL = 7
};

__global__ void k_gather_by_id(const SceneDev* __restrict__ sc, int p, int c, uint32_t n, int what,
                               void* out) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const SceneDev& S = *sc;
  const uint32_t id = S.id[p][s];
  switch (what) {
    case G_POS: reinterpret_cast<float4*>(out)[id] = S.x[p][s]; break;
    case G_PRED: {
      const float4 q = S.xs[c][s];
      reinterpret_cast<float4*>(out)[id] = make_float4(q.x, q.y, q.z, 0.0f);
      break;
    }
    case G_VEL: reinterpret_cast<float4*>(out)[id] = S.v[p][s]; break;
    case G_INVMASS: reinterpret_cast<float*>(out)[id] = S.w[p][s]; break;
    case G_DENSITY: reinterpret_cast<float*>(out)[id] = S.rho[p][s]; break;
    case G_KEY: reinterpret_cast<uint32_t*>(out)[id] = S.tmp_key[s]; break;
    case G_IDS: reinterpret_cast<uint32_t*>(out)[s] = id; break;
    case G_BVOL: reinterpret_cast<float*>(out)[id] = S.bvol[p][s]; break;
  }
}

void launch_gather_by_id(hipStream_t st, const SceneDev* sc, int p, int c, uint32_t n, int what,
                         void* out) {
  if (!n) return;
  hipLaunchKernelGGL(k_gather_by_id, dim3((n + 255) / 256), dim3(256), 0, st, sc, p, c, n, what, out);
}

// mask bits: 1 pos (x and x*), 2 vel, 4 inv_mass, 8 group, 16 density (value in inv_mass array)
__global__ void k_scatter_upload(const SceneDev* __restrict__ sc, int p, int c, uint32_t start,
                                 uint32_t count, const float4* pos, const float4* vel,
                                 const float* inv_mass, const uint32_t* group, uint32_t mask) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  const SceneDev& S = *sc;
  const uint32_t id = start + i;
  const uint32_t s = S.slot_of[id];
  if (mask & 1u) {
    S.x[p][s] = pos[i];
    S.xs[c][s] = pos[i];
  }
  if (mask & 2u) S.v[p][s] = vel[i];
  if (mask & 4u) S.w[p][s] = inv_mass[i];
  if (mask & 8u) S.grp[p][s] = group[i];
  if (mask & 16u) S.rho[p][s] = inv_mass[i];
  if (mask & 32u) S.bvol[p][s] = inv_mass[i];
}

void launch_scatter_upload(hipStream_t st, const SceneDev* sc, int p, int c, uint32_t start,
                           uint32_t count, const float4* pos, const float4* vel, const float* inv_mass,
                           const uint32_t* group, uint32_t mask) {
  if (!count) return;
  hipLaunchKernelGGL(k_scatter_upload, dim3((count + 255) / 256), dim3(256), 0, st, sc, p, c, start,
                     count, pos, vel, inv_mass, group, mask);
}

__global__ void k_iota(uint32_t* a, uint32_t* b, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    a[i] = i;
    b[i] = i;
  }
}

// Candidate count of the 27-bucket walk around each query point, excluding
// the query's own index (emitForEachNeighbor, query.ts:24-70), and the
// number of those within `radius` (the hash-grid tests' kernels).
__global__ void k_grid_query(const SceneDev* __restrict__ sc, const uint32_t* __restrict__ perm,
                             const float4* __restrict__ spos, const float4* __restrict__ q,
                             uint32_t nq, float r2, uint32_t* cand, uint32_t* within) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nq) return;
  const SceneDev& S = *sc;
  const float4 x = q[i];
  const int qx = pf_query_coord(x.x, S.ox, S.cell);
  const int qy = pf_query_coord(x.y, S.oy, S.cell);
  const int qz = pf_query_coord(x.z, S.oz, S.cell);
  uint32_t visited[27];
  uint32_t nv = 0, nc = 0, nw = 0;
  for (int dz = -1; dz <= 1; ++dz)
    for (int dy = -1; dy <= 1; ++dy)
      for (int dx = -1; dx <= 1; ++dx) {
        const uint32_t b = pf_morton((int)((uint32_t)qx + (uint32_t)dx), (int)((uint32_t)qy + (uint32_t)dy),
                                     (int)((uint32_t)qz + (uint32_t)dz)) & S.hash_mask;
        bool seen = false;
        for (uint32_t k = 0; k < nv; ++k) seen |= visited[k] == b;
        visited[nv++] = b;
        if (seen) continue;
        for (uint32_t j = S.start[b]; j < S.start[b + 1]; ++j) {
          if (perm[j] == i) continue;
          ++nc;
          const float4 y = spos[j];
          const float ddx = x.x - y.x, ddy = x.y - y.y, ddz = x.z - y.z;
          if ((ddx * ddx + ddy * ddy) + ddz * ddz < r2) ++nw;
        }
      }
  cand[i] = nc;
  within[i] = nw;
}

}  // namespace pf

using namespace pf;

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

#define PF_REQUIRE(ptr, owner)                                               \
  do {                                                                       \
    if (!(ptr)) fail(PF_EINVAL, "%s: null handle", owner);                   \
  } while (0)

extern "C" {

const char* pf_last_error(void) { return pf::g_error.c_str(); }
uint32_t pf_abi_version(void) { return 1u; }

// ---------------------------------------------------------------------------
// Context
// ---------------------------------------------------------------------------
int pf_context_create(int device, pf_context** out) {
  PF_API_BEGIN
  PF_REQUIRE(out, "pf_context_create");
  int ndev = 0;
  PF_HIP(hipGetDeviceCount(&ndev));
  if (device < 0 || device >= ndev) fail(PF_EINVAL, "pf_context_create: device %d out of range (%d devices)", device, ndev);
  PF_HIP(hipSetDevice(device));
  auto* ctx = new pf_context_s();
  ctx->device = device;
  PF_HIP(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
  PF_HIP(hipEventCreate(&ctx->t0));
  PF_HIP(hipEventCreate(&ctx->t1));
  ctx->scan_ctl = dalloc<uint32_t>(4);
  *out = ctx;
  PF_API_END
}

int pf_context_destroy(pf_context* ctx) {
  PF_API_BEGIN
  if (!ctx) return PF_OK;
  (void)hipStreamSynchronize(ctx->stream);
  for (auto& p : ctx->pending) {
    (void)hipEventDestroy(p.second.first);
    (void)hipEventDestroy(p.second.second);
  }
  for (auto e : ctx->event_pool) (void)hipEventDestroy(e);
  (void)hipEventDestroy(ctx->t0);
  (void)hipEventDestroy(ctx->t1);
  dfree(ctx->scan_ctl);
  (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  PF_API_END
}

int pf_context_synchronize(pf_context* ctx) {
  PF_API_BEGIN
  PF_REQUIRE(ctx, "pf_context_synchronize");
  PF_HIP(hipStreamSynchronize(ctx->stream));
  PF_API_END
}

int pf_context_timer_begin(pf_context* ctx) {
  PF_API_BEGIN
  PF_REQUIRE(ctx, "pf_context_timer_begin");
  PF_HIP(hipEventRecord(ctx->t0, ctx->stream));
  PF_API_END
}

int pf_context_timer_end(pf_context* ctx, float* ms) {
  PF_API_BEGIN
  PF_REQUIRE(ctx, "pf_context_timer_end");
  PF_HIP(hipEventRecord(ctx->t1, ctx->stream));
  PF_HIP(hipEventSynchronize(ctx->t1));
  float v = 0.f;
  PF_HIP(hipEventElapsedTime(&v, ctx->t0, ctx->t1));
  if (ms) *ms = v;
  PF_API_END
}

int pf_context_set_profiling(pf_context* ctx, int on) {
  PF_API_BEGIN
  PF_REQUIRE(ctx, "pf_context_set_profiling");
  if (!on) ctx->collect();
  ctx->profiling = on != 0;
  PF_API_END
}

int pf_context_profile_count(pf_context* ctx, uint32_t* n) {
  PF_API_BEGIN
  PF_REQUIRE(ctx, "pf_context_profile_count");
  ctx->collect();
  *n = (uint32_t)ctx->stats.size();
  PF_API_END
}

int pf_context_profile_entry(pf_context* ctx, uint32_t index, char* name, uint32_t name_cap,
                             uint32_t* launches, double* total_ms) {
  PF_API_BEGIN
  PF_REQUIRE(ctx, "pf_context_profile_entry");
  ctx->collect();
  if (index >= ctx->stats.size()) fail(PF_EINVAL, "pf_context_profile_entry: index %u out of range", index);
  auto it = ctx->stats.begin();
  std::advance(it, index);
  if (name && name_cap) {
    std::snprintf(name, name_cap, "%s", it->first.c_str());
  }
  if (launches) *launches = it->second.launches;
  if (total_ms) *total_ms = it->second.total_ms;
  PF_API_END
}

int pf_context_profile_reset(pf_context* ctx) {
  PF_API_BEGIN
  PF_REQUIRE(ctx, "pf_context_profile_reset");
  ctx->collect();
  ctx->stats.clear();
  PF_API_END
}

// ---------------------------------------------------------------------------
// ParticleSystem
// ---------------------------------------------------------------------------
static pf_particles* particles_new(pf_context* ctx, uint32_t capacity, double radius, bool evictable) {
  PF_REQUIRE(ctx, "ParticleSystem");
  if (capacity == 0) fail(PF_EINVAL, "ParticleSystem: capacity must be a positive integer, got %u", capacity);
  if (!(radius > 0) || !std::isfinite(radius))
    fail(PF_EINVAL, "ParticleSystem: particleRadius must be a positive finite number, got %g", radius);
  std::unique_ptr<pf_particles_s> owner(new pf_particles_s());
  pf_particles_s* ps = owner.get();
  // The arena's first segment is sized for the scene a fluid of this many
  // particles builds (state, sort, neighbour lists, tiles, contact records:
  // about 1 KB each).
  if (evictable) ps->arena.reset(new Arena(ctx->device, (size_t)capacity * kArenaBytesPerParticle));
  ArenaScope scope(ps->arena.get());
  ps->ctx = ctx;
  ps->n = capacity;
  ps->radius = radius;
  SceneDev& S = ps->host;
  S.n = capacity;
  S.radius = (float)radius;
  for (int k = 0; k < 2; ++k) {
    S.x[k] = dalloc<float4>(capacity);
    S.v[k] = dalloc<float4>(capacity);
    S.w[k] = dalloc<float>(capacity);
    S.id[k] = dalloc<uint32_t>(capacity);
    S.rho[k] = dalloc<float>(capacity);
    S.bvol[k] = dalloc<float>(capacity);
    S.grp[k] = dalloc<uint32_t>(capacity);
    S.xs[k] = dalloc<float4>(capacity);
  }
  S.mass = dalloc<float>(capacity);
  S.slot_of = dalloc<uint32_t>(capacity);
  S.flags = dalloc<uint32_t>(2 * PF_NUM_FLAGS);
  hipLaunchKernelGGL(k_iota, dim3((capacity + 255) / 256), dim3(256), 0, ctx->stream, S.id[0], S.slot_of, capacity);  PF_HIP(hipGetLastError());
  ps->dev = dalloc<SceneDev>(4);
  ps->cpu_inv_mass.assign(capacity, 0.0f);
  ps->cpu_positions.assign(3 * (size_t)capacity, 0.0f);
  ps->cpu_groups.assign(capacity, 0u);
  ps->dirty = true;
  ps->upload_desc();
  return owner.release();
}

int pf_particles_create(pf_context* ctx, uint32_t capacity, double radius, pf_particles** out) {
  PF_API_BEGIN
  PF_REQUIRE(out, "ParticleSystem");
  *out = particles_new(ctx, capacity, radius, false);
  PF_API_END
}

int pf_particles_create_evictable(pf_context* ctx, uint32_t capacity, double radius, pf_particles** out) {
  PF_API_BEGIN
  PF_REQUIRE(out, "ParticleSystem");
  *out = particles_new(ctx, capacity, radius, true);
  PF_API_END
}

int pf_particles_device_bytes(pf_particles* ps, uint64_t* bytes) {
  PF_API_BEGIN
  PF_REQUIRE(ps, "ParticleSystem");
  PF_REQUIRE(bytes, "ParticleSystem");
  *bytes = ps->arena ? (uint64_t)ps->arena->bytes() : 0;
  PF_API_END
}

static void particles_free(pf_particles* ps) {
  SceneDev& S = ps->host;
  for (int k = 0; k < 2; ++k) {
    dfree(S.x[k]); dfree(S.v[k]); dfree(S.w[k]); dfree(S.id[k]); dfree(S.rho[k]);
    dfree(S.bvol[k]); dfree(S.grp[k]); dfree(S.xs[k]);
  }
  dfree(S.mass); dfree(S.slot_of); dfree(S.flags); dfree(ps->dev);
  ps->dev = nullptr;
  ps->arena.reset();  // every other buffer of an evictable scene goes with it
}

int pf_particles_destroy(pf_particles* ps) {
  PF_API_BEGIN
  if (!ps || ps->disposed) return PF_OK;
  if (ps->server) fail(PF_ESTATE, "ParticleSystem: remove the scene from its server before disposing it");
  (void)hipStreamSynchronize(ps->ctx->stream);
  ps->disposed = true;
  particles_free(ps);
  PF_API_END
}

static void check_alive(pf_particles* ps, const char* owner) {
  if (!ps) fail(PF_EINVAL, "%s: null handle", owner);
  if (ps->disposed) fail(PF_ESTATE, "ParticleSystem has been disposed");
  if (ps->server) fail(PF_ESTATE, "%s: the scene is being served; remove it from the server first", owner);
}

static void check_range(pf_particles* ps, uint32_t start, uint32_t count, const char* owner) {
  if ((uint64_t)start + count > ps->n)
    fail(PF_EINVAL, "%s: invalid particle range start=%u count=%u (capacity %u)", owner, start, count, ps->n);
}

}  // extern "C"

namespace pf {
// Scatter host arrays (original order) into the cell-sorted device state.
void particles_scatter(pf_particles* ps, uint32_t start, uint32_t count, const std::vector<float4>* pos,
                       const std::vector<float4>* vel, const std::vector<float>* f,
                       const std::vector<uint32_t>* grp, uint32_t mask) {
  if (count == 0) return;
  ps->upload_desc();
  Context* ctx = ps->ctx;
  float4* dpos = nullptr; float4* dvel = nullptr; float* df = nullptr; uint32_t* dg = nullptr;
  if (pos) { dpos = dalloc_temp<float4>(count); PF_HIP(hipMemcpy(dpos, pos->data(), count * sizeof(float4), hipMemcpyHostToDevice)); }
  if (vel) { dvel = dalloc_temp<float4>(count); PF_HIP(hipMemcpy(dvel, vel->data(), count * sizeof(float4), hipMemcpyHostToDevice)); }
  if (f) { df = dalloc_temp<float>(count); PF_HIP(hipMemcpy(df, f->data(), count * sizeof(float), hipMemcpyHostToDevice)); }
  if (grp) { dg = dalloc_temp<uint32_t>(count); PF_HIP(hipMemcpy(dg, grp->data(), count * sizeof(uint32_t), hipMemcpyHostToDevice)); }
  launch_scatter_upload(ctx->stream, ps->cur(), 0, 0, start, count, dpos, dvel, df, dg, mask);
  PF_HIP(hipGetLastError());
  PF_HIP(hipStreamSynchronize(ctx->stream));
  dfree(dpos); dfree(dvel); dfree(df); dfree(dg);
}

// Gather one field into original order on the host.
void particles_gather(pf_particles* ps, int what, void* host, size_t elem) {
  ps->upload_desc();
  Context* ctx = ps->ctx;
  void* d = nullptr;
  d = dalloc_bytes((size_t)ps->n * elem, true);
  launch_gather_by_id(ctx->stream, ps->cur(), 0, 0, ps->n, what, d);
  PF_HIP(hipGetLastError());
  PF_HIP(hipMemcpyAsync(host, d, (size_t)ps->n * elem, hipMemcpyDeviceToHost, ctx->stream));
  PF_HIP(hipStreamSynchronize(ctx->stream));
  dfree(d);
}
}  // namespace pf

extern "C" {

int pf_particles_upload(pf_particles* ps, uint32_t start, uint32_t count, const float* positions,
                        const float* velocities, const float* inv_mass, const uint32_t* group) {
  PF_API_BEGIN
  check_alive(ps, "ParticleSystem.uploadParticles");
  if (count == 0) {
    if (start > ps->n)
      fail(PF_EINVAL, "ParticleSystem.uploadParticles: invalid particle range start=%u count=0 (capacity %u)", start, ps->n);
    return PF_OK;
  }
  check_range(ps, start, count, "ParticleSystem.uploadParticles");
  if (!positions) fail(PF_EINVAL, "ParticleSystem.uploadParticles: positions are required");
  std::vector<float4> pos(count), vel(count);
  std::vector<float> im(count);
  std::vector<uint32_t> gr(count);
  for (uint32_t k = 0; k < count; ++k) {
    if (inv_mass && !(inv_mass[k] >= 0.0f && std::isfinite(inv_mass[k])))
      fail(PF_EINVAL, "ParticleSystem.uploadParticles: invMass must be finite and >= 0, got %g", (double)inv_mass[k]);
  }
  for (uint32_t k = 0; k < count; ++k) {
    pos[k] = make_float4(positions[3 * k], positions[3 * k + 1], positions[3 * k + 2], 0.0f);
    vel[k] = velocities ? make_float4(velocities[3 * k], velocities[3 * k + 1], velocities[3 * k + 2], 0.0f)
                        : make_float4(0.f, 0.f, 0.f, 0.f);
    im[k] = inv_mass ? inv_mass[k] : 1.0f;
    gr[k] = group ? group[k] : 0u;
    ps->cpu_inv_mass[start + k] = im[k];
    ps->cpu_groups[start + k] = gr[k];
    for (int a = 0; a < 3; ++a) ps->cpu_positions[3 * (size_t)(start + k) + a] = positions[3 * k + a];
  }
  particles_scatter(ps, start, count, &pos, &vel, &im, &gr, 1u | 2u | 4u | 8u);
  PF_API_END
}

int pf_particles_set_inv_mass(pf_particles* ps, uint32_t start, uint32_t count, double inv_mass) {
  PF_API_BEGIN
  check_alive(ps, "ParticleSystem.setInvMass");
  check_range(ps, start, count, "ParticleSystem.setInvMass");
  if (!(inv_mass >= 0) || !std::isfinite(inv_mass))
    fail(PF_EINVAL, "ParticleSystem.setInvMass: invMass must be finite and >= 0, got %g", inv_mass);
  std::vector<float> im(count, (float)inv_mass);
  for (uint32_t k = 0; k < count; ++k) ps->cpu_inv_mass[start + k] = (float)inv_mass;
  particles_scatter(ps, start, count, nullptr, nullptr, &im, nullptr, 4u);
  PF_API_END
}

int pf_particles_set_collision_group(pf_particles* ps, uint32_t start, uint32_t count, uint32_t group) {
  PF_API_BEGIN
  check_alive(ps, "ParticleSystem.setCollisionGroup");
  check_range(ps, start, count, "ParticleSystem.setCollisionGroup");
  std::vector<uint32_t> g(count, group);
  for (uint32_t k = 0; k < count; ++k) ps->cpu_groups[start + k] = group;
  particles_scatter(ps, start, count, nullptr, nullptr, nullptr, &g, 8u);
  PF_API_END
}

int pf_particles_readback(pf_particles* ps, const pf_snapshot* out) {
  PF_API_BEGIN
  check_alive(ps, "ParticleSystem.readback");
  PF_REQUIRE(out, "ParticleSystem.readback");
  if (ps->loop) PF_HIP(hipStreamSynchronize(ps->ctx->stream));
  if (out->positions) particles_gather(ps, G_POS, out->positions, sizeof(float4));
  if (out->predicted_positions) particles_gather(ps, G_PRED, out->predicted_positions, sizeof(float4));
  if (out->velocities) particles_gather(ps, G_VEL, out->velocities, sizeof(float4));
  if (out->inv_mass) particles_gather(ps, G_INVMASS, out->inv_mass, sizeof(float));
  // Orientations only change under local shape matching, whose loop keeps
  // them in particle order; otherwise every particle stays at the identity.
  const pf_softbody_s* sb = ps->loop && ps->loop->soft && ps->loop->soft->mode == PF_SHAPE_MATCHING_LOCAL
                                ? ps->loop->soft : nullptr;
  const size_t bytes = sizeof(float) * 4 * (size_t)ps->n;
  if (sb) {
    if (out->rotation) PF_HIP(hipMemcpy(out->rotation, sb->d_q, bytes, hipMemcpyDeviceToHost));
    if (out->predicted_rotation) PF_HIP(hipMemcpy(out->predicted_rotation, sb->d_qp, bytes, hipMemcpyDeviceToHost));
    if (out->angular_velocity) PF_HIP(hipMemcpy(out->angular_velocity, sb->d_omega, bytes, hipMemcpyDeviceToHost));
  } else {
    for (float* q : {out->rotation, out->predicted_rotation}) {
      if (!q) continue;
      for (uint32_t i = 0; i < ps->n; ++i) { q[4 * i] = 0; q[4 * i + 1] = 0; q[4 * i + 2] = 0; q[4 * i + 3] = 1; }
    }
    if (out->angular_velocity) std::memset(out->angular_velocity, 0, bytes);
  }
  PF_API_END
}

int pf_particles_readback_boundary_volume(pf_particles* ps, float* volume) {
  PF_API_BEGIN
  check_alive(ps, "ParticleSystem.readback");
  PF_REQUIRE(volume, "ParticleSystem.readback");
  if (ps->loop) PF_HIP(hipStreamSynchronize(ps->ctx->stream));
  particles_gather(ps, G_BVOL, volume, sizeof(float));
  PF_API_END
}

int pf_particles_info(pf_particles* ps, uint32_t* capacity, double* radius) {
  PF_API_BEGIN
  check_alive(ps, "ParticleSystem");
  if (capacity) *capacity = ps->n;
  if (radius) *radius = ps->radius;
  PF_API_END
}

// ---------------------------------------------------------------------------
// HashGrid
// ---------------------------------------------------------------------------
static uint32_t next_pow2(uint64_t n) {
  if (n <= 1) return 1;
  uint64_t p = 1;
  while (p < n) p <<= 1;
  return (uint32_t)std::min<uint64_t>(p, 1ull << 31);
}

static void grid_init(pf_grid* g, pf::Context* ctx, uint32_t n, double cell, uint32_t H,
                      const double origin[3], const char* owner) {
  if (!(cell > 0) || !std::isfinite(cell))
    fail(PF_EINVAL, "HashGrid: cellSize must be a positive finite number, got %g", cell);
  if (H == 0) H = std::min<uint32_t>(next_pow2(2ull * n), 1u << 20);
  if ((H & (H - 1)) != 0 || H > (1u << 30))
    fail(PF_EINVAL, "HashGrid: hashTableSize must be a positive power of two, got %u", H);
  (void)owner;
  g->ctx = ctx;
  g->n = n;
  g->H = H;
  g->Hpad = (uint32_t)(((uint64_t)H + 1023) / 1024 * 1024);
  g->cell = cell;
  if (origin) for (int a = 0; a < 3; ++a) g->origin[a] = origin[a];
  SceneDev& S = g->host;
  S.n = n;
  S.H = H;
  S.hash_mask = H - 1;
  S.cell = (float)cell;
  S.ox = (float)g->origin[0];
  S.oy = (float)g->origin[1];
  S.oz = (float)g->origin[2];
  g->positions = dalloc<float4>(n);
  S.x[0] = g->positions;
  S.key = dalloc<uint32_t>(n);
  S.start = dalloc<uint32_t>((size_t)H + 1);
  S.tmp_src = dalloc<uint32_t>(n);
  S.tmp_key = dalloc<uint32_t>(n);
  S.flags = dalloc<uint32_t>(2 * PF_NUM_FLAGS);
  g->perm = dalloc<uint32_t>(n);
  g->skey = dalloc<uint32_t>(n);
  g->spos = dalloc<float4>(n);
  g->radix = dalloc<uint32_t>(grid_radix_scratch_words(n, H));
  g->dev = dalloc<SceneDev>(1);
  PF_HIP(hipMemcpy(g->dev, &S, sizeof(SceneDev), hipMemcpyHostToDevice));
}

static void grid_free(pf_grid* g) {
  if (g->owned_by_loop) return;
  SceneDev& S = g->host;
  dfree(g->positions); dfree(S.key); dfree(S.start); dfree(S.tmp_src); dfree(S.tmp_key); dfree(S.flags);
  dfree(g->perm); dfree(g->skey); dfree(g->spos); dfree(g->radix); dfree(g->dev);
}

int pf_grid_create(pf_particles* ps, double cell_size, uint32_t hash_table_size,
                   const double origin[3], pf_grid** out) {
  PF_API_BEGIN
  check_alive(ps, "HashGrid");
  PF_REQUIRE(out, "HashGrid");
  ArenaScope arena_scope(ps->arena.get());
  auto* g = new pf_grid_s();
  try {
    grid_init(g, ps->ctx, ps->n, cell_size, hash_table_size, origin, "HashGrid");
  } catch (...) {
    delete g;
    throw;
  }
  g->ps = ps;
  *out = g;
  PF_API_END
}

int pf_grid_create_standalone(pf_context* ctx, uint32_t capacity, double cell_size,
                              uint32_t hash_table_size, const double origin[3], pf_grid** out) {
  PF_API_BEGIN
  PF_REQUIRE(ctx, "HashGrid");
  PF_REQUIRE(out, "HashGrid");
  if (capacity == 0) fail(PF_EINVAL, "HashGrid: capacity must be a positive integer, got %u", capacity);
  auto* g = new pf_grid_s();
  try {
    grid_init(g, ctx, capacity, cell_size, hash_table_size, origin, "HashGrid");
  } catch (...) {
    delete g;
    throw;
  }
  *out = g;
  PF_API_END
}

static void grid_alive(pf_grid* g, const char* what) {
  if (!g) fail(PF_EINVAL, "%s: null handle", what);
  if (g->disposed) fail(PF_ESTATE, "HashGrid has been disposed");
}

int pf_grid_upload_positions(pf_grid* g, const float* positions4, uint32_t count) {
  PF_API_BEGIN
  grid_alive(g, "HashGrid.uploadPositions");
  if (g->ps || g->owned_by_loop) fail(PF_EINVAL, "HashGrid: positions come from its ParticleSystem");
  if (count > g->n) fail(PF_EINVAL, "HashGrid: %u positions exceed capacity %u", count, g->n);
  PF_HIP(hipStreamSynchronize(g->ctx->stream));
  PF_HIP(hipMemcpy(g->positions, positions4, (size_t)count * sizeof(float4), hipMemcpyHostToDevice));
  PF_API_END
}

int pf_grid_set_sorted_positions(pf_grid* g, int enabled) {
  PF_API_BEGIN
  grid_alive(g, "HashGrid");
  g->sorted_positions = enabled != 0;
  PF_API_END
}

int pf_grid_rebuild(pf_grid* g) {
  PF_API_BEGIN
  grid_alive(g, "HashGrid.rebuild");
  if (g->owned_by_loop) fail(PF_EINVAL, "HashGrid: the SimLoop rebuilds its own grid every substep");
  Context* ctx = g->ctx;
  hipStream_t st = ctx->stream;
  if (g->ps) {
    check_alive(g->ps, "HashGrid.rebuild");
    g->ps->upload_desc();
    PF_LAUNCH(ctx, "gather_predicted");
    launch_gather_by_id(st, g->ps->cur(), 0, 0, g->n, G_PRED, g->positions);
  }
  {
    // The overflow flag covers one rebuild (HashGrid.ts:197, cellIndex.ts:109-111).
    PF_HIP(hipMemsetAsync(g->host.flags, 0, sizeof(uint32_t) * 2 * PF_NUM_FLAGS, st));
  }
  {
    // keys, radix passes and bucket offsets, timed as one sort
    PF_LAUNCH(ctx, "grid_sort");
    launch_grid_radix(st, g->dev, g->host, g->skey, g->perm, g->sorted_positions ? g->spos : nullptr, g->radix);
  }
  PF_HIP(hipGetLastError());
  g->built = true;
  PF_API_END
}

int pf_grid_readback(pf_grid* g, uint32_t* cell_index, uint32_t* counts, uint32_t* cell_start,
                     uint32_t* cell_end, uint32_t* sorted_indices) {
  PF_API_BEGIN
  grid_alive(g, "HashGrid.readback");
  Context* ctx = g->ctx;
  hipStream_t st = ctx->stream;
  PF_HIP(hipStreamSynchronize(st));
  const SceneDev& S = g->host;
  std::vector<uint32_t> start((size_t)g->H + 1);
  PF_HIP(hipMemcpy(start.data(), S.start, start.size() * sizeof(uint32_t), hipMemcpyDeviceToHost));
  const uint32_t total = start[g->H];
  for (uint32_t b = 0; b < g->Hpad; ++b) {
    const uint32_t s0 = b < g->H ? start[b] : total;
    const uint32_t s1 = b < g->H ? start[b + 1] : total;
    if (counts) counts[b] = s1 - s0;
    if (cell_start) cell_start[b] = s0;
    if (cell_end) cell_end[b] = s1;
  }
  if (g->owned_by_loop) {
    pf_particles* ps = g->ps;
    if (cell_index) particles_gather(ps, G_KEY, cell_index, sizeof(uint32_t));
    if (sorted_indices) {
      void* d = nullptr;
      d = dalloc_bytes((size_t)g->n * 4, true);
      launch_gather_by_id(st, ps->cur(), 0, 0, g->n, G_IDS, d);
      PF_HIP(hipMemcpyAsync(sorted_indices, d, (size_t)g->n * 4, hipMemcpyDeviceToHost, st));
      PF_HIP(hipStreamSynchronize(st));
      dfree(d);
    }
  } else {
    if (cell_index) PF_HIP(hipMemcpy(cell_index, S.key, (size_t)g->n * 4, hipMemcpyDeviceToHost));
    if (sorted_indices) PF_HIP(hipMemcpy(sorted_indices, g->perm, (size_t)g->n * 4, hipMemcpyDeviceToHost));
  }
  PF_API_END
}

int pf_grid_readback_sorted_keys(pf_grid* g, uint32_t* keys) {
  PF_API_BEGIN
  grid_alive(g, "HashGrid.readback");
  PF_HIP(hipStreamSynchronize(g->ctx->stream));
  PF_HIP(hipMemcpy(keys, g->owned_by_loop ? g->host.tmp_key : g->skey, (size_t)g->n * 4, hipMemcpyDeviceToHost));
  PF_API_END
}

int pf_grid_readback_sorted_positions(pf_grid* g, float* positions4) {
  PF_API_BEGIN
  grid_alive(g, "HashGrid.readback");
  if (g->owned_by_loop) fail(PF_EINVAL, "HashGrid: a SimLoop grid keeps positions in the particle state");
  if (!g->sorted_positions) fail(PF_EINVAL, "HashGrid: sorted positions are disabled");
  PF_HIP(hipStreamSynchronize(g->ctx->stream));
  PF_HIP(hipMemcpy(positions4, g->spos, (size_t)g->n * sizeof(float4), hipMemcpyDeviceToHost));
  PF_API_END
}

int pf_grid_readback_overflow(pf_grid* g, int* overflow) {
  PF_API_BEGIN
  grid_alive(g, "HashGrid.readbackOverflow");
  PF_HIP(hipStreamSynchronize(g->ctx->stream));
  uint32_t f[2 * PF_NUM_FLAGS];
  PF_HIP(hipMemcpy(f, g->host.flags, sizeof f, hipMemcpyDeviceToHost));
  const uint32_t half = g->owned_by_loop && g->ps->loop ? g->ps->loop->last_flag_half : 0u;
  *overflow = f[half * PF_NUM_FLAGS + PF_FLAG_GRID] != 0;
  PF_API_END
}

int pf_grid_info(pf_grid* g, uint32_t* capacity, uint32_t* H, uint32_t* Hpad) {
  PF_API_BEGIN
  grid_alive(g, "HashGrid");
  if (capacity) *capacity = g->n;
  if (H) *H = g->H;
  if (Hpad) *Hpad = g->Hpad;
  PF_API_END
}

int pf_grid_set_origin(pf_grid* g, const double origin[3]) {
  PF_API_BEGIN
  grid_alive(g, "HashGrid");
  PF_REQUIRE(origin, "HashGrid");
  for (int a = 0; a < 3; ++a) g->origin[a] = origin[a];
  if (g->owned_by_loop) {
    pf_particles* ps = g->ps;
    ps->host.ox = (float)origin[0]; ps->host.oy = (float)origin[1]; ps->host.oz = (float)origin[2];
    ps->dirty = true;
  } else {
    g->host.ox = (float)origin[0]; g->host.oy = (float)origin[1]; g->host.oz = (float)origin[2];
    PF_HIP(hipStreamSynchronize(g->ctx->stream));
    PF_HIP(hipMemcpy(g->dev, &g->host, sizeof(SceneDev), hipMemcpyHostToDevice));
  }
  PF_API_END
}

int pf_grid_destroy(pf_grid* g) {
  PF_API_BEGIN
  if (!g || g->disposed) return PF_OK;
  if (g->owned_by_loop) return PF_OK;  // freed with its loop
  (void)hipStreamSynchronize(g->ctx->stream);
  g->disposed = true;
  grid_free(g);
  PF_API_END
}

// Neighbour-walk candidate counts over a built standalone grid (test support).
int pf_grid_query_counts(pf_grid* g, const float* query4, uint32_t nq, double radius,
                         uint32_t* candidates, uint32_t* within) {
  PF_API_BEGIN
  grid_alive(g, "HashGrid.query");
  if (g->owned_by_loop || !g->sorted_positions) fail(PF_EINVAL, "HashGrid.query: needs a standalone grid with sorted positions");
  hipStream_t st = g->ctx->stream;
  float4* dq = dalloc_temp<float4>(nq);
  uint32_t* dc = dalloc_temp<uint32_t>(nq);
  uint32_t* dw = dalloc_temp<uint32_t>(nq);
  PF_HIP(hipMemcpy(dq, query4, (size_t)nq * sizeof(float4), hipMemcpyHostToDevice));
  hipLaunchKernelGGL(k_grid_query, dim3((nq + 255) / 256), dim3(256), 0, st, g->dev, g->perm, g->spos, dq, nq,
                     (float)(radius * radius), dc, dw);
  PF_HIP(hipGetLastError());
  PF_HIP(hipStreamSynchronize(st));
  if (candidates) PF_HIP(hipMemcpy(candidates, dc, (size_t)nq * 4, hipMemcpyDeviceToHost));
  if (within) PF_HIP(hipMemcpy(within, dw, (size_t)nq * 4, hipMemcpyDeviceToHost));
  dfree(dq); dfree(dc); dfree(dw);
  PF_API_END
}

}  // extern "C"
