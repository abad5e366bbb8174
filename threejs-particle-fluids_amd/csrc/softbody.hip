// Global XPBD shape matching for gfx950 (shapeMatch.ts:53-351,
// polarDecomp.ts:34-264, SoftbodySystem.ts:254-275).
//
// The reference runs three dispatches per solver iteration (centre of mass,
// moment + polar decomposition, delta apply), each one workgroup per body
// with an LDS tree. Here one kernel does all three per body: the strided
// per-thread partial sums, the tree fold (LDS for the two cross-wave levels,
// wave64 shuffles below 64, pairing lane i with i + s exactly as the
// reference's tree so the FP32 sums are identical), the 3x3 polar
// decomposition on one lane, then the goal pull. Bodies are independent, so
// no grid-wide barrier is needed between the reduction and the pull.
//
// Body particles are addressed by original id; `slot_of` maps them to their
// cell-sorted slot.
#include "pf_internal.hpp"
#include "pf_kernels.hpp"

namespace pf {

struct Mat3 {
  float m[3][3];
};

// polarDecomp.ts:34-264 (8 sweeps of cyclic Jacobi on A^T A, then R = A S^-1).
__device__ Mat3 polar_decomposition(const Mat3& a) {
  float m00 = (a.m[0][0] * a.m[0][0] + a.m[1][0] * a.m[1][0]) + a.m[2][0] * a.m[2][0];
  float m01 = (a.m[0][0] * a.m[0][1] + a.m[1][0] * a.m[1][1]) + a.m[2][0] * a.m[2][1];
  float m02 = (a.m[0][0] * a.m[0][2] + a.m[1][0] * a.m[1][2]) + a.m[2][0] * a.m[2][2];
  float m11 = (a.m[0][1] * a.m[0][1] + a.m[1][1] * a.m[1][1]) + a.m[2][1] * a.m[2][1];
  float m12 = (a.m[0][1] * a.m[0][2] + a.m[1][1] * a.m[1][2]) + a.m[2][1] * a.m[2][2];
  float m22 = (a.m[0][2] * a.m[0][2] + a.m[1][2] * a.m[1][2]) + a.m[2][2] * a.m[2][2];
  float q00 = 1.f, q01 = 0.f, q02 = 0.f, q10 = 0.f, q11 = 1.f, q12 = 0.f, q20 = 0.f, q21 = 0.f, q22 = 1.f;
  const float eps = 1e-12f;
  for (int sweep = 0; sweep < 8; ++sweep) {
    // plane (0,1)
    if (fabsf(m01) > eps) {
      const float apq = m01, app = m00, aqq = m11;
      const float tau = (aqq - app) / (apq * 2.0f);
      const float sg = tau >= 0.0f ? 1.0f : -1.0f;
      const float ta = fabsf(tau);
      const float t = sg / (ta + sqrtf(ta * ta + 1.0f));
      const float cc = 1.0f / sqrtf(t * t + 1.0f);
      const float ss = t * cc;
      const float o02 = m02, o12 = m12;
      const float a00 = q00, a01 = q01, a10 = q10, a11 = q11, a20 = q20, a21 = q21;
      m00 = app - t * apq;
      m11 = aqq + t * apq;
      m01 = 0.0f;
      m02 = cc * o02 - ss * o12;
      m12 = ss * o02 + cc * o12;
      q00 = cc * a00 - ss * a01; q01 = ss * a00 + cc * a01;
      q10 = cc * a10 - ss * a11; q11 = ss * a10 + cc * a11;
      q20 = cc * a20 - ss * a21; q21 = ss * a20 + cc * a21;
    }
    // plane (0,2)
    if (fabsf(m02) > eps) {
      const float apq = m02, app = m00, aqq = m22;
      const float tau = (aqq - app) / (apq * 2.0f);
      const float sg = tau >= 0.0f ? 1.0f : -1.0f;
      const float ta = fabsf(tau);
      const float t = sg / (ta + sqrtf(ta * ta + 1.0f));
      const float cc = 1.0f / sqrtf(t * t + 1.0f);
      const float ss = t * cc;
      const float o01 = m01, o12 = m12;
      const float a00 = q00, a02 = q02, a10 = q10, a12 = q12, a20 = q20, a22 = q22;
      m00 = app - t * apq;
      m22 = aqq + t * apq;
      m02 = 0.0f;
      m01 = cc * o01 - ss * o12;
      m12 = ss * o01 + cc * o12;
      q00 = cc * a00 - ss * a02; q02 = ss * a00 + cc * a02;
      q10 = cc * a10 - ss * a12; q12 = ss * a10 + cc * a12;
      q20 = cc * a20 - ss * a22; q22 = ss * a20 + cc * a22;
    }
    // plane (1,2)
    if (fabsf(m12) > eps) {
      const float apq = m12, app = m11, aqq = m22;
      const float tau = (aqq - app) / (apq * 2.0f);
      const float sg = tau >= 0.0f ? 1.0f : -1.0f;
      const float ta = fabsf(tau);
      const float t = sg / (ta + sqrtf(ta * ta + 1.0f));
      const float cc = 1.0f / sqrtf(t * t + 1.0f);
      const float ss = t * cc;
      const float o01 = m01, o02 = m02;
      const float a01 = q01, a02 = q02, a11 = q11, a12 = q12, a21 = q21, a22 = q22;
      m11 = app - t * apq;
      m22 = aqq + t * apq;
      m12 = 0.0f;
      m01 = cc * o01 - ss * o02;
      m02 = ss * o01 + cc * o02;
      q01 = cc * a01 - ss * a02; q02 = ss * a01 + cc * a02;
      q11 = cc * a11 - ss * a12; q12 = ss * a11 + cc * a12;
      q21 = cc * a21 - ss * a22; q22 = ss * a21 + cc * a22;
    }
  }
  const float i0 = 1.0f / sqrtf(fmaxf(m00, 1e-20f));
  const float i1 = 1.0f / sqrtf(fmaxf(m11, 1e-20f));
  const float i2 = 1.0f / sqrtf(fmaxf(m22, 1e-20f));
  const float s00 = ((q00 * q00) * i0 + (q01 * q01) * i1) + (q02 * q02) * i2;
  const float s01 = ((q00 * q10) * i0 + (q01 * q11) * i1) + (q02 * q12) * i2;
  const float s02 = ((q00 * q20) * i0 + (q01 * q21) * i1) + (q02 * q22) * i2;
  const float s11 = ((q10 * q10) * i0 + (q11 * q11) * i1) + (q12 * q12) * i2;
  const float s12 = ((q10 * q20) * i0 + (q11 * q21) * i1) + (q12 * q22) * i2;
  const float s22 = ((q20 * q20) * i0 + (q21 * q21) * i1) + (q22 * q22) * i2;
  const float si[3][3] = {{s00, s01, s02}, {s01, s11, s12}, {s02, s12, s22}};
  Mat3 r;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      r.m[i][j] = (a.m[i][0] * si[0][j] + a.m[i][1] * si[1][j]) + a.m[i][2] * si[2][j];
  return r;
}

// Fold 256 per-thread values with the reference tree pairing (s = 128 .. 1).
template <int N>
__device__ __forceinline__ void tree_fold(float (&v)[N], float (*lds)[PF_BLOCK]) {
  const uint32_t tid = threadIdx.x;
#pragma unroll
  for (int k = 0; k < N; ++k) lds[k][tid] = v[k];
  __syncthreads();
#pragma unroll
  for (uint32_t s = PF_BLOCK / 2; s >= 64; s >>= 1) {
    if (tid < s) {
#pragma unroll
      for (int k = 0; k < N; ++k) lds[k][tid] = lds[k][tid] + lds[k][tid + s];
    }
    __syncthreads();
  }
  if (tid < 64) {
#pragma unroll
    for (int k = 0; k < N; ++k) v[k] = lds[k][tid];
#pragma unroll
    for (int s = 32; s > 0; s >>= 1) {
#pragma unroll
      for (int k = 0; k < N; ++k) v[k] = v[k] + __shfl_down(v[k], s);
    }
  }
}

__global__ __launch_bounds__(PF_BLOCK) void k_shape_match(const SceneDev* __restrict__ scenes,
                                                          const uint2* __restrict__ bodies,
                                                          int p, int c, int first_iteration,
                                                          int pull) {
  __shared__ float lds[13][PF_BLOCK];
  __shared__ float sR[9], sC[3];
  const uint2 bt = bodies[blockIdx.x];
  const SceneDev& S = scenes[bt.x];
  const SoftDev& B = S.soft;
  const uint32_t b = bt.y;
  const uint32_t start = B.body_start[b];
  const uint32_t count = B.body_count[b];
  const uint32_t tid = threadIdx.x;
  float4* __restrict__ xs = S.xs[c];
  // Strided partial sums: centre (x m, m) and moment rows r (m x)_row.
  float acc[13];
#pragma unroll
  for (int k = 0; k < 13; ++k) acc[k] = 0.0f;
  for (uint32_t i = tid; i < count; i += PF_BLOCK) {
    const uint32_t id = start + i;
    const uint32_t slot = S.slot_of[id];
    const float4 x = xs[slot];
    const float m = B.weights[id];
    acc[0] = acc[0] + x.x * m;
    acc[1] = acc[1] + x.y * m;
    acc[2] = acc[2] + x.z * m;
    acc[3] = acc[3] + m;
    const float xmx = x.x * m, xmy = x.y * m, xmz = x.z * m;
    const float4 r = B.rest[id];
    acc[4] = acc[4] + r.x * xmx; acc[5] = acc[5] + r.y * xmx; acc[6] = acc[6] + r.z * xmx;
    acc[7] = acc[7] + r.x * xmy; acc[8] = acc[8] + r.y * xmy; acc[9] = acc[9] + r.z * xmy;
    acc[10] = acc[10] + r.x * xmz; acc[11] = acc[11] + r.y * xmz; acc[12] = acc[12] + r.z * xmz;
  }
  tree_fold<13>(acc, lds);
  if (tid == 0) {
    const float cx = acc[0] / acc[3], cy = acc[1] / acc[3], cz = acc[2] / acc[3];
    Mat3 a;
    a.m[0][0] = acc[4]; a.m[0][1] = acc[5]; a.m[0][2] = acc[6];
    a.m[1][0] = acc[7]; a.m[1][1] = acc[8]; a.m[1][2] = acc[9];
    a.m[2][0] = acc[10]; a.m[2][1] = acc[11]; a.m[2][2] = acc[12];
    const Mat3 R = polar_decomposition(a);
    B.centers[b] = make_float4(cx, cy, cz, 0.0f);
    for (int r = 0; r < 3; ++r) B.rotations[3 * b + r] = make_float4(R.m[r][0], R.m[r][1], R.m[r][2], 0.0f);
    sC[0] = cx; sC[1] = cy; sC[2] = cz;
    for (int k = 0; k < 9; ++k) sR[k] = R.m[k / 3][k % 3];
  }
  __syncthreads();
  if (!pull) return;
  const float at = B.body_compliance[b] / (S.dt * S.dt);
  const float R00 = sR[0], R01 = sR[1], R02 = sR[2], R10 = sR[3], R11 = sR[4], R12 = sR[5],
              R20 = sR[6], R21 = sR[7], R22 = sR[8];
  const float cx = sC[0], cy = sC[1], cz = sC[2];
  for (uint32_t i = tid; i < count; i += PF_BLOCK) {
    const uint32_t id = start + i;
    const uint32_t slot = S.slot_of[id];
    const float4 r = B.rest[id];
    const float gx = ((R00 * r.x + R01 * r.y) + R02 * r.z) + cx;
    const float gy = ((R10 * r.x + R11 * r.y) + R12 * r.z) + cy;
    const float gz = ((R20 * r.x + R21 * r.y) + R22 * r.z) + cz;
    const float4 x = xs[slot];
    const float Cx = x.x - gx, Cy = x.y - gy, Cz = x.z - gz;
    const float4 lo = first_iteration ? make_float4(0.f, 0.f, 0.f, 0.f) : B.lambda[id];
    const float w = S.w[p][slot];
    const float den = w + at;
    const float dlx = ((-Cx) - lo.x * at) / den;
    const float dly = ((-Cy) - lo.y * at) / den;
    const float dlz = ((-Cz) - lo.z * at) / den;
    B.lambda[id] = make_float4(lo.x + dlx, lo.y + dly, lo.z + dlz, 0.0f);
    xs[slot] = make_float4(x.x + dlx * w, x.y + dly * w, x.z + dlz * w, 0.0f);
  }
}

void launch_shape_match(hipStream_t st, const SceneDev* sc, const uint2* bodies, uint32_t nbodies,
                        int p, int c, int first_iteration, int pull) {
  if (!nbodies) return;
  hipLaunchKernelGGL(k_shape_match, dim3(nbodies), dim3(PF_BLOCK), 0, st, sc, bodies, p, c,
                     first_iteration, pull);
}

}  // namespace pf
