// Exhaustive device check of the correctly rounded reciprocal and square root
// the passes use (pf_rcp_rn, pf_sqrt_rn in pf_internal.hpp), and of the
// one-instruction saturating conversion (pf_f2i_sat_hw): each is compared bit
// for bit with its reference form (1.0f / x, sqrtf, pf_f2i_sat) over a range
// of float bit patterns, on the GPU that runs the passes. They rest on the
// hardware's own instructions, so this is where their exactness is shown.
#include <hip/hip_runtime.h>

#include "runtime.hpp"

namespace pf {
namespace {

constexpr uint32_t kFirst = 8;  // mismatching bit patterns kept

__global__ __launch_bounds__(PF_BLOCK) void k_exact_math(int function, uint32_t lo, uint32_t hi,
                                                         unsigned long long* count, uint32_t* first) {
  const uint64_t stride = (uint64_t)gridDim.x * PF_BLOCK;
  uint32_t local = 0;
  for (uint64_t b = (uint64_t)lo + (uint64_t)blockIdx.x * PF_BLOCK + threadIdx.x; b < hi; b += stride) {
    const float x = __uint_as_float((uint32_t)b);
    float ref, got;
    if (function == PF_EXACT_F2I) {
      ref = __int_as_float(pf_f2i_sat(x));
      got = __int_as_float(pf_f2i_sat_hw(x));
    } else if (function == PF_EXACT_RCP) {
      ref = 1.0f / x;
      got = pf_rcp_rn(x);
    } else if (function == PF_EXACT_SQRT) {
      ref = sqrtf(x);
      got = pf_sqrt_rn(x);
    } else {
      ref = sqrtf(x);
      got = pf_sqrt_rn_fast(x);
    }
    if (__float_as_uint(ref) != __float_as_uint(got)) {
      ++local;
      const unsigned long long k = atomicAdd(count + 1, 1ull);
      if (k < kFirst) first[k] = (uint32_t)b;
    }
  }
  if (local) atomicAdd(count, (unsigned long long)local);
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

int pf_check_exact_math(pf_context* ctx, int function, uint32_t first_bits, uint32_t end_bits,
                        uint64_t* mismatches, uint32_t* first_mismatch) {
  using namespace pf;
  PF_API_BEGIN
  if (!ctx) fail(PF_EINVAL, "pf_check_exact_math: null handle");
  if (function != PF_EXACT_RCP && function != PF_EXACT_SQRT && function != PF_EXACT_SQRT_FAST &&
      function != PF_EXACT_F2I)
    fail(PF_EINVAL, "pf_check_exact_math: unknown function %d", function);
  if (!mismatches) fail(PF_EINVAL, "pf_check_exact_math: mismatches is required");
  if (end_bits < first_bits) fail(PF_EINVAL, "pf_check_exact_math: empty range");
  int prev = 0;
  PF_HIP(hipGetDevice(&prev));
  if (prev != ctx->device) PF_HIP(hipSetDevice(ctx->device));
  struct Restore {
    int dev, prev;
    ~Restore() {
      if (dev != prev) (void)hipSetDevice(prev);
    }
  } restore{ctx->device, prev};
  unsigned long long* d_count = nullptr;
  uint32_t* d_first = nullptr;
  PF_HIP(hipMalloc(&d_count, 2 * sizeof(unsigned long long)));
  struct Free {
    void* a;
    void* b;
    ~Free() {
      (void)hipFree(a);
      (void)hipFree(b);
    }
  } fr{d_count, nullptr};
  PF_HIP(hipMalloc(&d_first, kFirst * sizeof(uint32_t)));
  fr.b = d_first;
  hipStream_t st = ctx->stream;
  PF_HIP(hipMemsetAsync(d_count, 0, 2 * sizeof(unsigned long long), st));
  PF_HIP(hipMemsetAsync(d_first, 0xff, kFirst * sizeof(uint32_t), st));
  hipLaunchKernelGGL(k_exact_math, dim3(8192), dim3(PF_BLOCK), 0, st, function, first_bits, end_bits, d_count,
                     d_first);
  PF_HIP(hipGetLastError());
  unsigned long long h_count[2] = {0, 0};
  uint32_t h_first = 0xffffffffu;
  PF_HIP(hipMemcpyAsync(h_count, d_count, sizeof(h_count), hipMemcpyDeviceToHost, st));
  PF_HIP(hipMemcpyAsync(&h_first, d_first, sizeof(h_first), hipMemcpyDeviceToHost, st));
  PF_HIP(hipStreamSynchronize(st));
  *mismatches = h_count[0];
  if (first_mismatch) *first_mismatch = h_first;
  PF_API_END
}

}  // extern "C"
