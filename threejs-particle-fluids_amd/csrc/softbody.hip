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
  // pull 0: refresh the body frames of local-matching scenes after the
  // substep (their solve does not fit body frames).
  if (!pull && !B.local) return;
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
  if (B.local) return;  // local shape matching pulls through its own kernels
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

// ---------------------------------------------------------------------------
// Local shape matching with oriented particles (shapeMatchImplicit.ts,
// rotation.ts): every particle's group is itself plus its edge neighbours.
// Kernels run over particle ids of the soft-body span; positions are read
// and written through slot_of.
// ---------------------------------------------------------------------------

// One (scene, first id) tile of the span; false past its end or when the
// scene does not use local matching.
static __device__ __forceinline__ bool local_id(const SceneDev* scenes, const uint2* tiles,
                                                const SceneDev** Sp, uint32_t* id) {
  const uint2 t = tiles[blockIdx.x];
  const SceneDev& S = scenes[t.x];
  *Sp = &S;
  *id = t.y + threadIdx.x;
  return S.soft.local && *id < S.soft.span_hi;
}

// Orientation predict (rotation.ts:38-55): qp = rotate(omega, dt) * q.
__global__ __launch_bounds__(PF_BLOCK) void k_orient_predict(const SceneDev* __restrict__ scenes,
                                                             const uint2* __restrict__ tiles) {
  const SceneDev* Sp;
  uint32_t i;
  if (!local_id(scenes, tiles, &Sp, &i)) return;
  const SceneDev& S = *Sp;
  const SoftDev& B = S.soft;
  const float4 q = B.q[i];
  const float3 om = pf_xyz(B.omega[i]);
  const float mag = pf_len3(om);
  if (mag < 1e-8f) {
    B.qp[i] = q;
    return;
  }
  const float half = (mag * S.dt) * 0.5f;
  const float sr = sinf(half) / mag;
  const float3 sp = make_float3(om.x * sr, om.y * sr, om.z * sr);
  const float sw = cosf(half);
  const float3 qv = pf_xyz(q);
  const float3 nv = pf_add3(pf_add3(pf_mul3(qv, sw), pf_mul3(sp, q.w)), pf_cross3(sp, qv));
  B.qp[i] = make_float4(nv.x, nv.y, nv.z, sw * q.w - pf_dot3(sp, qv));
}

// Step 1 (shapeMatchImplicit.ts:47-71): group centre c_i, the mean of the
// members' predicted positions.
__global__ __launch_bounds__(PF_BLOCK) void k_local_center(const SceneDev* __restrict__ scenes,
                                                           const uint2* __restrict__ tiles, int c) {
  const SceneDev* Sp;
  uint32_t i;
  if (!local_id(scenes, tiles, &Sp, &i)) return;
  const SceneDev& S = *Sp;
  const SoftDev& B = S.soft;
  const uint32_t o0 = B.csr_off[i], o1 = B.csr_off[i + 1];
  float3 sum = make_float3(0.f, 0.f, 0.f);
  for (uint32_t k = o0; k < o1; ++k) sum = pf_add3(sum, pf_xyz(S.xs[c][S.slot_of[B.csr_idx[k]]]));
  const uint32_t cnt = o1 - o0;
  const float inv = 1.0f / (float)(cnt > 1u ? cnt : 1u);
  B.group_center[i] = cnt == 0u ? make_float4(0.f, 0.f, 0.f, 0.f)
                                : make_float4(sum.x * inv, sum.y * inv, sum.z * inv, 0.0f);
}

// Step 2 (shapeMatchImplicit.ts:100-217): A_i = (r^2/5) sum_j R(qp_j)
// + sum_j (x*_j - c_i)(rest_j - rest centre_i)^T, R_i = polar(A_i).
__global__ __launch_bounds__(PF_BLOCK) void k_local_rotation(const SceneDev* __restrict__ scenes,
                                                             const uint2* __restrict__ tiles, int c) {
  const SceneDev* Sp;
  uint32_t i;
  if (!local_id(scenes, tiles, &Sp, &i)) return;
  const SceneDev& S = *Sp;
  const SoftDev& B = S.soft;
  const uint32_t o0 = B.csr_off[i], o1 = B.csr_off[i + 1];
  float a[3][3] = {{0.f, 0.f, 0.f}, {0.f, 0.f, 0.f}, {0.f, 0.f, 0.f}};
  const float3 ci = pf_xyz(B.group_center[i]);
  const float3 cb = pf_xyz(B.rest_center[i]);
  const float ai = B.ai;
  for (uint32_t k = o0; k < o1; ++k) {
    const uint32_t j = B.csr_idx[k];
    const float4 qj = B.qp[j];
    const float xx = qj.x * qj.x, yy = qj.y * qj.y, zz = qj.z * qj.z;
    const float xy = qj.x * qj.y, xz = qj.x * qj.z, yz = qj.y * qj.z;
    const float wx = qj.w * qj.x, wy = qj.w * qj.y, wz = qj.w * qj.z;
    const float rj[3][3] = {{1.0f - (yy + zz) * 2.0f, (xy - wz) * 2.0f, (xz + wy) * 2.0f},
                            {(xy + wz) * 2.0f, 1.0f - (xx + zz) * 2.0f, (yz - wx) * 2.0f},
                            {(xz - wy) * 2.0f, (yz + wx) * 2.0f, 1.0f - (xx + yy) * 2.0f}};
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
      for (int cc = 0; cc < 3; ++cc) a[r][cc] = a[r][cc] + ai * rj[r][cc];
    const float3 xj = pf_sub3(pf_xyz(S.xs[c][S.slot_of[j]]), ci);
    const float3 tj = pf_sub3(pf_xyz(B.rest[j]), cb);
    const float xv[3] = {xj.x, xj.y, xj.z}, tv[3] = {tj.x, tj.y, tj.z};
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
      for (int cc = 0; cc < 3; ++cc) a[r][cc] = a[r][cc] + xv[r] * tv[cc];
  }
  Mat3 A;
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int cc = 0; cc < 3; ++cc) A.m[r][cc] = a[r][cc];
  Mat3 R = polar_decomposition(A);
  if (o1 == o0) {
    for (int r = 0; r < 3; ++r)
      for (int cc = 0; cc < 3; ++cc) R.m[r][cc] = r == cc ? 1.0f : 0.0f;
  }
  for (int r = 0; r < 3; ++r) B.group_rot[3 * (size_t)i + r] = make_float4(R.m[r][0], R.m[r][1], R.m[r][2], 0.0f);
}

// Step 3 (shapeMatchImplicit.ts:248-308): pull every member toward
// R_i (rest_j - rest centre_i) + c_i with its own multiplier, the correction
// shared among the member's groups and clamped to 5 r, into the accumulator.
__global__ __launch_bounds__(PF_BLOCK) void k_local_scatter(const SceneDev* __restrict__ scenes,
                                                            const uint2* __restrict__ tiles, int p, int c,
                                                            int first_iteration) {
  const SceneDev* Sp;
  uint32_t i;
  if (!local_id(scenes, tiles, &Sp, &i)) return;
  const SceneDev& S = *Sp;
  const SoftDev& B = S.soft;
  const uint32_t o0 = B.csr_off[i], o1 = B.csr_off[i + 1];
  const float3 R0 = pf_xyz(B.group_rot[3 * (size_t)i]);
  const float3 R1 = pf_xyz(B.group_rot[3 * (size_t)i + 1]);
  const float3 R2 = pf_xyz(B.group_rot[3 * (size_t)i + 2]);
  const float3 ci = pf_xyz(B.group_center[i]);
  const float3 cb = pf_xyz(B.rest_center[i]);
  const float at = B.rest[i].w / (S.dt * S.dt);
  for (uint32_t k = o0; k < o1; ++k) {
    const uint32_t j = B.csr_idx[k];
    const uint32_t slot = S.slot_of[j];
    const float3 rr = pf_sub3(pf_xyz(B.rest[j]), cb);
    const float3 goal = pf_add3(make_float3(pf_dot3(R0, rr), pf_dot3(R1, rr), pf_dot3(R2, rr)), ci);
    const float3 C = pf_sub3(pf_xyz(S.xs[c][slot]), goal);
    const float wj = S.w[p][slot];
    const float den = wj + at;
    const float4 lo = first_iteration ? make_float4(0.f, 0.f, 0.f, 0.f) : B.pair_lambda[k];
    const float3 dl = make_float3(((-C.x) - lo.x * at) / den, ((-C.y) - lo.y * at) / den, ((-C.z) - lo.z * at) / den);
    B.pair_lambda[k] = make_float4(lo.x + dl.x, lo.y + dl.y, lo.z + dl.z, 0.0f);
    const uint32_t mcnt = B.csr_off[j + 1] - B.csr_off[j];
    const float mj = fmaxf((float)mcnt, 1.0f);
    const float3 dx = make_float3((dl.x * wj) / mj, (dl.y * wj) / mj, (dl.z * wj) / mj);
    const float sc = sqrtf(B.maxc2 / fmaxf(pf_dot3(dx, dx), B.maxc2));
    atomicAdd(B.acc + 3 * (size_t)j + 0, pf_ticks(dx.x * sc, S.pos_scale));
    atomicAdd(B.acc + 3 * (size_t)j + 1, pf_ticks(dx.y * sc, S.pos_scale));
    atomicAdd(B.acc + 3 * (size_t)j + 2, pf_ticks(dx.z * sc, S.pos_scale));
  }
}

// The pulls' accumulator apply over the span (accumulator.ts:101-131), then
// step 4 (shapeMatchImplicit.ts:325-402): R_i as the predicted orientation,
// on the shorter arc from the substep-start orientation.
__global__ __launch_bounds__(PF_BLOCK) void k_local_apply(const SceneDev* __restrict__ scenes,
                                                          const uint2* __restrict__ tiles, int c) {
  const SceneDev* Sp;
  uint32_t i;
  if (!local_id(scenes, tiles, &Sp, &i)) return;
  const SceneDev& S = *Sp;
  const SoftDev& B = S.soft;
  int* acc = B.acc + 3 * (size_t)i;
  const int tk[3] = {acc[0], acc[1], acc[2]};
  const float tx = (float)tk[0], ty = (float)tk[1], tz = (float)tk[2];
  const uint32_t slot = S.slot_of[i];
  const float4 x = S.xs[c][slot];
  S.xs[c][slot] = make_float4(x.x + tx * S.pos_inv, x.y + ty * S.pos_inv, x.z + tz * S.pos_inv, x.w);
  acc[0] = 0;
  acc[1] = 0;
  acc[2] = 0;
}

__global__ __launch_bounds__(PF_BLOCK) void k_local_orient(const SceneDev* __restrict__ scenes,
                                                           const uint2* __restrict__ tiles) {
  const SceneDev* Sp;
  uint32_t i;
  if (!local_id(scenes, tiles, &Sp, &i)) return;
  const SoftDev& B = Sp->soft;
  if (!(B.csr_off[i + 1] > B.csr_off[i])) return;
  const float4 r0 = B.group_rot[3 * (size_t)i], r1 = B.group_rot[3 * (size_t)i + 1], r2 = B.group_rot[3 * (size_t)i + 2];
  const float m00 = r0.x, m01 = r0.y, m02 = r0.z, m10 = r1.x, m11 = r1.y, m12 = r1.z, m20 = r2.x, m21 = r2.y, m22 = r2.z;
  const float trace = (m00 + m11) + m22;
  float qx, qy, qz, qw;
  if (trace > 0.0f) {
    const float s = 0.5f / sqrtf(trace + 1.0f);
    qw = 0.25f / s;
    qx = (m21 - m12) * s;
    qy = (m02 - m20) * s;
    qz = (m10 - m01) * s;
  } else if (m00 > m11 && m00 > m22) {
    const float s = 2.0f * sqrtf(((1.0f + m00) - m11) - m22);
    qw = (m21 - m12) / s;
    qx = 0.25f * s;
    qy = (m01 + m10) / s;
    qz = (m02 + m20) / s;
  } else if (m11 > m22) {
    const float s = 2.0f * sqrtf(((1.0f + m11) - m00) - m22);
    qw = (m02 - m20) / s;
    qx = (m01 + m10) / s;
    qy = 0.25f * s;
    qz = (m12 + m21) / s;
  } else {
    const float s = 2.0f * sqrtf(((1.0f + m22) - m00) - m11);
    qw = (m10 - m01) / s;
    qx = (m02 + m20) / s;
    qy = (m12 + m21) / s;
    qz = 0.25f * s;
  }
  const float4 qs = B.q[i];
  const float d = ((qs.x * qx + qs.y * qy) + qs.z * qz) + qs.w * qw;
  const float f = d < 0.0f ? -1.0f : 1.0f;
  B.qp[i] = make_float4(qx * f, qy * f, qz * f, qw * f);
}

// Orientation advect (rotation.ts:57-82): omega from qp * q^-1 on the
// shorter arc, capped at 100 rad/s, then q = normalised qp.
__global__ __launch_bounds__(PF_BLOCK) void k_orient_advect(const SceneDev* __restrict__ scenes,
                                                            const uint2* __restrict__ tiles) {
  const SceneDev* Sp;
  uint32_t i;
  if (!local_id(scenes, tiles, &Sp, &i)) return;
  const SceneDev& S = *Sp;
  const SoftDev& B = S.soft;
  const float4 q = B.q[i];
  const float4 raw = B.qp[i];
  const float l = sqrtf(fmaxf(((raw.x * raw.x + raw.y * raw.y) + raw.z * raw.z) + raw.w * raw.w, 1e-12f));
  const float4 qp = make_float4(raw.x / l, raw.y / l, raw.z / l, raw.w / l);
  const float3 qi = make_float3(-q.x, -q.y, -q.z);
  const float3 pv = pf_xyz(qp);
  const float3 rpre = pf_add3(pf_add3(pf_mul3(pv, q.w), pf_mul3(qi, qp.w)), pf_cross3(pv, qi));
  const float wpre = qp.w * q.w - pf_dot3(pv, qi);
  const float f = wpre < 0.0f ? -1.0f : 1.0f;
  const float3 rv = pf_mul3(rpre, f);
  const float rw = wpre * f;
  const float hs = pf_len3(rv);
  if (hs < 1e-8f) {
    B.omega[i] = make_float4(0.f, 0.f, 0.f, 0.f);
  } else {
    const float angle = atan2f(hs, rw) * 2.0f;
    const float3 om = pf_mul3(rv, (angle / hs) / S.dt);
    const float sc = sqrtf(10000.0f / fmaxf(pf_dot3(om, om), 10000.0f));
    B.omega[i] = make_float4(om.x * sc, om.y * sc, om.z * sc, 0.0f);
  }
  B.q[i] = qp;
}

void launch_local_predict(hipStream_t st, const SceneDev* sc, const uint2* tiles, uint32_t ntiles) {
  if (!ntiles) return;
  hipLaunchKernelGGL(k_orient_predict, dim3(ntiles), dim3(PF_BLOCK), 0, st, sc, tiles);
}

void launch_local_solve(hipStream_t st, const SceneDev* sc, const uint2* tiles, uint32_t ntiles, int p, int c,
                        int first_iteration) {
  if (!ntiles) return;
  hipLaunchKernelGGL(k_local_center, dim3(ntiles), dim3(PF_BLOCK), 0, st, sc, tiles, c);
  hipLaunchKernelGGL(k_local_rotation, dim3(ntiles), dim3(PF_BLOCK), 0, st, sc, tiles, c);
  hipLaunchKernelGGL(k_local_scatter, dim3(ntiles), dim3(PF_BLOCK), 0, st, sc, tiles, p, c, first_iteration);
  hipLaunchKernelGGL(k_local_apply, dim3(ntiles), dim3(PF_BLOCK), 0, st, sc, tiles, c);
  hipLaunchKernelGGL(k_local_orient, dim3(ntiles), dim3(PF_BLOCK), 0, st, sc, tiles);
}

void launch_local_advect(hipStream_t st, const SceneDev* sc, const uint2* tiles, uint32_t ntiles) {
  if (!ntiles) return;
  hipLaunchKernelGGL(k_orient_advect, dim3(ntiles), dim3(PF_BLOCK), 0, st, sc, tiles);
}

void launch_shape_match(hipStream_t st, const SceneDev* sc, const uint2* bodies, uint32_t nbodies,
                        int p, int c, int first_iteration, int pull) {
  if (!nbodies) return;
  hipLaunchKernelGGL(k_shape_match, dim3(nbodies), dim3(PF_BLOCK), 0, st, sc, bodies, p, c,
                     first_iteration, pull);
}

}  // namespace pf
