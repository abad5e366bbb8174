// Analytic collider primitives and XPBD contact/friction for one particle
// (primitives.ts:85-242, collider.ts:53-118, PrimitiveSet.ts:328-405).
// Device-only, included by the kernels that fuse collider work.
#pragma once
#include "pf_internal.hpp"

namespace pf {

// v + 2w(q x v) + q x (2(q x v)) (primitives.ts:22-27).
__device__ __forceinline__ float3 quat_rotate(float4 q, float3 v) {
  const float3 qv = make_float3(q.x, q.y, q.z);
  const float3 c = pf_cross3(qv, v);
  const float3 t = make_float3(2.0f * c.x, 2.0f * c.y, 2.0f * c.z);
  return pf_add3(pf_add3(v, pf_mul3(t, q.w)), pf_cross3(qv, t));
}

__device__ __forceinline__ float4 quat_conj(float4 q) { return make_float4(-q.x, -q.y, -q.z, q.w); }

// primitives.ts:33-38
__device__ __forceinline__ float4 quat_from_spin(float3 spin, float t) {
  const float rate = pf_len3(spin);
  const float half = (rate * t) * 0.5f;
  const float den = fmaxf(rate, 1e-8f);
  const float3 axis = make_float3(spin.x / den, spin.y / den, spin.z / den);
  const float s = sinf(half);
  return make_float4(axis.x * s, axis.y * s, axis.z * s, cosf(half));
}

__device__ __forceinline__ float3 prim_pivot(uint32_t kind, float4 d0, float4 d1) {
  if (kind == PF_KIND_PLANE) return pf_xyz(d1);
  if (kind == PF_KIND_CAPSULE) return pf_mul3(pf_add3(pf_xyz(d0), pf_xyz(d1)), 0.5f);
  return pf_xyz(d0);
}

// Signed distance and outward gradient of primitive `slot` at xw, evaluated
// as the primitive was `rewind` seconds ago when it sweeps (primitives.ts:85-217).
// Signed distance and gradient of primitive `slot` at x as it stands now
// (primitives.ts:119-217), inversion included.
__device__ __forceinline__ void prim_shape(const SceneDev& S, uint32_t slot, uint32_t kind, uint32_t flags,
                                           float4 d0, float4 d1, float3 x, float* phi_out, float3* grad_out);

__device__ __forceinline__ void prim_sdf(const SceneDev& S, uint32_t slot, float3 xw, float rewind,
                                         float* phi_out, float3* grad_out) {
  const uint32_t packed = S.prim_packed[slot];
  const uint32_t kind = (packed >> 16) & 0xffffu;
  const uint32_t flags = packed & 0xffffu;
  const float4 d0 = S.prim_d0[slot];
  const float4 d1 = S.prim_d1[slot];
  const float t = (flags & PF_PRIM_SWEEP) ? rewind : 0.0f;
  const float3 shifted = pf_add3(xw, pf_mul3(pf_xyz(S.prim_vel[slot]), t));
  const float3 pivot = prim_pivot(kind, d0, d1);
  const float4 turn = quat_from_spin(pf_xyz(S.prim_spin[slot]), t);
  const float3 x = pf_add3(pivot, quat_rotate(turn, pf_sub3(shifted, pivot)));
  float phi;
  float3 g;
  prim_shape(S, slot, kind, flags, d0, d1, x, &phi, &g);
  *phi_out = phi;
  *grad_out = quat_rotate(quat_conj(turn), g);
}

// The primitive's distance at xw with no sweep rewind (primitives.ts:85-92
// called without `rewind`).
__device__ __forceinline__ float prim_phi_now(const SceneDev& S, uint32_t slot, float3 xw) {
  const uint32_t packed = S.prim_packed[slot];
  float phi;
  float3 g;
  prim_shape(S, slot, (packed >> 16) & 0xffffu, packed & 0xffffu, S.prim_d0[slot], S.prim_d1[slot], xw, &phi, &g);
  return phi;
}

__device__ __forceinline__ void prim_shape(const SceneDev& S, uint32_t slot, uint32_t kind, uint32_t flags,
                                           float4 d0, float4 d1, float3 x, float* phi_out, float3* grad_out) {
  float phi = 0.0f;
  float3 g = make_float3(0.f, 0.f, 0.f);
  if (kind == PF_KIND_PLANE) {
    const float3 n = pf_xyz(d0);
    phi = pf_dot3(n, pf_sub3(x, pf_xyz(d1)));
    g = n;
  } else if (kind == PF_KIND_SPHERE) {
    const float3 diff = pf_sub3(x, pf_xyz(d0));
    const float len = pf_len3(diff);
    phi = len - d0.w;
    const float sl = fmaxf(len, 1e-8f);
    g = make_float3(diff.x / sl, diff.y / sl, diff.z / sl);
  } else if (kind == PF_KIND_BOX) {
    const float4 rot = S.prim_rot[slot];
    const float3 e = pf_xyz(d1);
    const float3 xl = quat_rotate(quat_conj(rot), pf_sub3(x, pf_xyz(d0)));
    const float3 q = make_float3(fabsf(xl.x) - e.x, fabsf(xl.y) - e.y, fabsf(xl.z) - e.z);
    const float3 qp = make_float3(fmaxf(q.x, 0.f), fmaxf(q.y, 0.f), fmaxf(q.z, 0.f));
    const float outside = pf_len3(qp);
    const float inside = fminf(fmaxf(fmaxf(q.x, q.y), q.z), 0.0f);
    phi = outside + inside;
    const float3 sg = make_float3(xl.x >= 0.f ? 1.f : -1.f, xl.y >= 0.f ? 1.f : -1.f,
                                  xl.z >= 0.f ? 1.f : -1.f);
    const bool is_out = outside > 1e-8f;
    const bool domx = q.x >= q.y && q.x >= q.z;
    const bool domy = q.y >= q.x && q.y >= q.z;
    float3 lg;
    if (is_out) {
      const float ol = fmaxf(outside, 1e-8f);
      lg = make_float3((qp.x * sg.x) / ol, (qp.y * sg.y) / ol, (qp.z * sg.z) / ol);
    } else {
      lg = make_float3(domx ? sg.x : 0.f, (!domx && domy) ? sg.y : 0.f,
                       (!domx && !domy) ? sg.z : 0.f);
    }
    g = quat_rotate(rot, lg);
  } else {  // capsule
    const float3 a = pf_xyz(d0);
    const float3 ab = pf_sub3(pf_xyz(d1), a);
    const float3 ax = pf_sub3(x, a);
    const float ll = fmaxf(pf_dot3(ab, ab), 1e-8f);
    const float tt = fminf(fmaxf(pf_dot3(ax, ab) / ll, 0.0f), 1.0f);
    const float3 closest = pf_add3(a, pf_mul3(ab, tt));
    const float3 diff = pf_sub3(x, closest);
    const float len = pf_len3(diff);
    phi = len - d0.w;
    const float sl = fmaxf(len, 1e-8f);
    g = make_float3(diff.x / sl, diff.y / sl, diff.z / sl);
  }
  if (flags & PF_PRIM_INVERT) {
    phi = -phi;
    g = make_float3(-g.x, -g.y, -g.z);
  }
  *phi_out = phi;
  *grad_out = g;
}

// v + omega x (x - c + v t) (primitives.ts:224-242).
__device__ __forceinline__ float3 prim_surface_velocity(const SceneDev& S, uint32_t slot, float3 xw,
                                                        float rewind) {
  const uint32_t packed = S.prim_packed[slot];
  const uint32_t kind = (packed >> 16) & 0xffffu;
  const float3 vel = pf_xyz(S.prim_vel[slot]);
  const float3 pivot = prim_pivot(kind, S.prim_d0[slot], S.prim_d1[slot]);
  const float t = (packed & PF_PRIM_SWEEP) ? rewind : 0.0f;
  const float3 arm = pf_add3(pf_sub3(xw, pivot), pf_mul3(vel, t));
  return pf_add3(vel, pf_cross3(pf_xyz(S.prim_spin[slot]), arm));
}

// One XPBD contact iteration against every primitive (PrimitiveSet.ts:336-364,
// collider.ts:53-89). Returns the position-accumulator ticks of this particle;
// (lambda_n, lambda_t) live per slot and primitive, and `touch` bit p marks
// that primitive p's multipliers hold this substep's values (first touch
// starts them at zero, as the per-substep reset does).
__device__ __forceinline__ void walls_solve(const SceneDev& S, uint32_t s, float3 xstar, float3 x,
                                            float w, float clock, bool first_iteration, int* ticks) {
  if (!(w > 0.0f) || S.nprim == 0) return;
  uint32_t touch = first_iteration ? 0u : S.prim_touch[s];
  const uint32_t touch_in = touch;
  const float3 disp = pf_sub3(xstar, x);
  for (uint32_t p = 0; p < S.nprim; ++p) {
    float phi;
    float3 n;
    prim_sdf(S, p, xstar, clock, &phi, &n);
    if (phi < S.radius) {
      const float3 cd = pf_mul3(prim_surface_velocity(S, p, xstar, clock), S.dt);
      const float depth = S.radius - phi;
      float* lam = S.prim_lambda + ((size_t)s * S.prim_cap + p) * 2u;
      const bool had = (touch >> p) & 1u;
      const float ln_old = had ? lam[0] : 0.0f;
      const float lt_old = had ? lam[1] : 0.0f;
      const float ln = ln_old + depth / w;
      const float3 slip = pf_sub3(disp, cd);
      const float3 tang = pf_sub3(slip, pf_mul3(n, pf_dot3(slip, n)));
      const float tl = pf_len3(tang);
      const float mus = S.prim_d1[p].w;
      const bool sticks = tl > 1e-10f && (tl / w) <= (mus * ln - lt_old);
      const float3 push = pf_mul3(n, depth);
      const float3 corr = sticks ? pf_sub3(push, tang) : push;
      ticks[0] = pf_wrap_add(ticks[0], pf_ticks(corr.x, S.pos_scale));
      ticks[1] = pf_wrap_add(ticks[1], pf_ticks(corr.y, S.pos_scale));
      ticks[2] = pf_wrap_add(ticks[2], pf_ticks(corr.z, S.pos_scale));
      lam[0] = ln;
      lam[1] = lt_old + (sticks ? tl / w : 0.0f);
      touch |= 1u << p;
    }
  }
  if (first_iteration || touch != touch_in) S.prim_touch[s] = touch;
}

// Kinetic friction against every primitive a particle pressed on this substep
// (PrimitiveSet.ts:366-391, collider.ts:96-118). Returns velocity ticks.
__device__ __forceinline__ void walls_friction(const SceneDev& S, uint32_t s, float3 xstar,
                                               float3 vel, float w, float clock, int* ticks) {
  if (!(w > 0.0f) || S.nprim == 0) return;
  const uint32_t touch = S.prim_touch[s];
  for (uint32_t p = 0; p < S.nprim; ++p) {
    if (!((touch >> p) & 1u)) continue;  // lambda_n is zero after the reset
    const float ln = S.prim_lambda[((size_t)s * S.prim_cap + p) * 2u];
    if (!(ln > 0.0f)) continue;
    float phi;
    float3 g;
    prim_sdf(S, p, xstar, clock, &phi, &g);
    const float gl = pf_len3(g);
    if (!(gl > 1e-6f)) continue;
    const float3 n = make_float3(g.x / gl, g.y / gl, g.z / gl);
    const float3 rel = pf_sub3(vel, prim_surface_velocity(S, p, xstar, clock));
    const float3 vt = pf_sub3(rel, pf_mul3(n, pf_dot3(n, rel)));
    const float vl = pf_len3(vt);
    if (vl > 1e-6f) {
      const float muk = S.prim_vel[p].w;
      const float change = fminf(((muk * ln) * w) / S.dt, vl);
      const float3 d = make_float3(-((vt.x / vl) * change), -((vt.y / vl) * change),
                                   -((vt.z / vl) * change));
      ticks[0] = pf_wrap_add(ticks[0], pf_ticks(d.x, S.vel_scale));
      ticks[1] = pf_wrap_add(ticks[1], pf_ticks(d.y, S.vel_scale));
      ticks[2] = pf_wrap_add(ticks[2], pf_ticks(d.z, S.vel_scale));
    }
  }
}

// ---------------------------------------------------------------------------
// SDF colliders (SDFCollider.ts:68-311, sdf.ts:107-146)
// ---------------------------------------------------------------------------
__device__ __forceinline__ float pf_half_to_float(uint16_t h) {
  return (float)__builtin_bit_cast(_Float16, h);
}

__device__ __forceinline__ float sdf_texel(const SdfDev& D, int i, int j, int k) {
  return pf_half_to_float(D.grid[((size_t)k * (uint32_t)D.ny + (uint32_t)j) * (uint32_t)D.nx + (uint32_t)i]);
}

__device__ __forceinline__ float sdf_lerp(float a, float b, float f) { return a + (b - a) * f; }

// Linear filtering at normalised uv with clamp-to-edge addressing: texel
// coordinate t = uv * n - 0.5, weights from its fraction, lerp x, y, z.
__device__ __forceinline__ float sdf_tex(const SdfDev& D, float u, float v, float w) {
  const float t[3] = {u * D.res[0] - 0.5f, v * D.res[1] - 0.5f, w * D.res[2] - 0.5f};
  const int n[3] = {D.nx, D.ny, D.nz};
  int i0[3], i1[3];
  float f[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const float fl = floorf(t[a]);
    f[a] = t[a] - fl;
    const int ii = pf_clampi(pf_f2i_sat(fl), -1, n[a] - 1);
    i0[a] = ii < 0 ? 0 : ii;
    i1[a] = ii + 1 > n[a] - 1 ? n[a] - 1 : ii + 1;
  }
  const float c00 = sdf_lerp(sdf_texel(D, i0[0], i0[1], i0[2]), sdf_texel(D, i1[0], i0[1], i0[2]), f[0]);
  const float c10 = sdf_lerp(sdf_texel(D, i0[0], i1[1], i0[2]), sdf_texel(D, i1[0], i1[1], i0[2]), f[0]);
  const float c01 = sdf_lerp(sdf_texel(D, i0[0], i0[1], i1[2]), sdf_texel(D, i1[0], i0[1], i1[2]), f[0]);
  const float c11 = sdf_lerp(sdf_texel(D, i0[0], i1[1], i1[2]), sdf_texel(D, i1[0], i1[1], i1[2]), f[0]);
  return sdf_lerp(sdf_lerp(c00, c10, f[1]), sdf_lerp(c01, c11, f[1]), f[2]);
}

// emitSampleSdf: world -> local -> uv, trilinear phi, 6-tap central-difference
// gradient, back to world.
__device__ __forceinline__ void sdf_sample(const SdfDev& D, float3 x, float* phi, float3* grad) {
  const float wo[3] = {x.x - D.pos[0], x.y - D.pos[1], x.z - D.pos[2]};
  float uv[3];
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    const float xl = ((D.inv_rot[3 * r] * wo[0] + D.inv_rot[3 * r + 1] * wo[1]) + D.inv_rot[3 * r + 2] * wo[2]) * D.inv_scale;
    uv[r] = ((xl - D.origin[r]) * D.inv_voxel[r]) / D.res[r];
  }
  const float pl = sdf_tex(D, uv[0], uv[1], uv[2]);
  float gl[3];
  gl[0] = (sdf_tex(D, uv[0] + D.inv_res[0], uv[1], uv[2]) - sdf_tex(D, uv[0] - D.inv_res[0], uv[1], uv[2])) / (2.0f * D.voxel[0]);
  gl[1] = (sdf_tex(D, uv[0], uv[1] + D.inv_res[1], uv[2]) - sdf_tex(D, uv[0], uv[1] - D.inv_res[1], uv[2])) / (2.0f * D.voxel[1]);
  gl[2] = (sdf_tex(D, uv[0], uv[1], uv[2] + D.inv_res[2]) - sdf_tex(D, uv[0], uv[1], uv[2] - D.inv_res[2])) / (2.0f * D.voxel[2]);
  *phi = pl * D.scale;
  *grad = make_float3((D.rot[0] * gl[0] + D.rot[1] * gl[1]) + D.rot[2] * gl[2],
                      (D.rot[3] * gl[0] + D.rot[4] * gl[1]) + D.rot[5] * gl[2],
                      (D.rot[6] * gl[0] + D.rot[7] * gl[1]) + D.rot[8] * gl[2]);
}

// One XPBD contact iteration against every SDF collider of the scene
// (SDFCollider.ts:237-269, collider.ts:53-89; static colliders). Ticks go to
// the positions accumulator; (lambda_n, lambda_t) per slot and collider with
// the same touch-mask reset as the primitives.
__device__ __forceinline__ void sdf_solve(const SceneDev& S, uint32_t s, float3 xstar, float3 x, float w,
                                          bool first_iteration, int* ticks) {
  if (!(w > 0.0f) || S.nsdf == 0) return;
  uint32_t touch = first_iteration ? 0u : S.sdf_touch[s];
  const uint32_t touch_in = touch;
  const float3 disp = pf_sub3(xstar, x);
  for (uint32_t k = 0; k < S.nsdf; ++k) {
    const SdfDev& D = S.sdf[k];
    float phi;
    float3 g;
    sdf_sample(D, xstar, &phi, &g);
    const float gl = pf_len3(g);
    if (!(gl > 1e-6f)) continue;
    const float3 n = make_float3(g.x / gl, g.y / gl, g.z / gl);
    const float reach = S.radius + D.thickness;
    if (!(phi < reach)) continue;
    const float depth = reach - phi;
    float* lam = S.sdf_lambda + ((size_t)s * PF_MAX_SDF + k) * 2u;
    const bool had = (touch >> k) & 1u;
    const float ln = (had ? lam[0] : 0.0f) + depth / w;
    const float lt_old = had ? lam[1] : 0.0f;
    const float3 tang = pf_sub3(disp, pf_mul3(n, pf_dot3(disp, n)));
    const float tl = pf_len3(tang);
    const bool sticks = tl > 1e-10f && (tl / w) <= (D.mus * ln - lt_old);
    const float3 push = pf_mul3(n, depth);
    const float3 corr = sticks ? pf_sub3(push, tang) : push;
    ticks[0] = pf_wrap_add(ticks[0], pf_ticks(corr.x, S.pos_scale));
    ticks[1] = pf_wrap_add(ticks[1], pf_ticks(corr.y, S.pos_scale));
    ticks[2] = pf_wrap_add(ticks[2], pf_ticks(corr.z, S.pos_scale));
    lam[0] = ln;
    lam[1] = lt_old + (sticks ? tl / w : 0.0f);
    touch |= 1u << k;
  }
  if (first_iteration || touch != touch_in) S.sdf_touch[s] = touch;
}

// Kinetic friction against every SDF collider the particle pressed on this
// substep (SDFCollider.ts:271-291, collider.ts:96-118). Velocity ticks.
__device__ __forceinline__ void sdf_friction(const SceneDev& S, uint32_t s, float3 xstar, float3 vel, float w,
                                             int* ticks) {
  if (!(w > 0.0f) || S.nsdf == 0) return;
  const uint32_t touch = S.sdf_touch[s];
  for (uint32_t k = 0; k < S.nsdf; ++k) {
    if (!((touch >> k) & 1u)) continue;
    const float ln = S.sdf_lambda[((size_t)s * PF_MAX_SDF + k) * 2u];
    if (!(ln > 0.0f)) continue;
    const SdfDev& D = S.sdf[k];
    float phi;
    float3 g;
    sdf_sample(D, xstar, &phi, &g);
    const float gl = pf_len3(g);
    if (!(gl > 1e-6f)) continue;
    const float3 n = make_float3(g.x / gl, g.y / gl, g.z / gl);
    const float3 vt = pf_sub3(vel, pf_mul3(n, pf_dot3(n, vel)));
    const float vl = pf_len3(vt);
    if (vl > 1e-6f) {
      const float change = fminf(((D.muk * ln) * w) / S.dt, vl);
      const float3 d = make_float3(-((vt.x / vl) * change), -((vt.y / vl) * change), -((vt.z / vl) * change));
      ticks[0] = pf_wrap_add(ticks[0], pf_ticks(d.x, S.vel_scale));
      ticks[1] = pf_wrap_add(ticks[1], pf_ticks(d.y, S.vel_scale));
      ticks[2] = pf_wrap_add(ticks[2], pf_ticks(d.z, S.vel_scale));
    }
  }
}

// Accumulator apply (accumulator.ts:101-131): ticks -> f32, saturation test,
// scaled change. Returns the change; sets *saturated.
__device__ __forceinline__ float3 ticks_to_change(const int* ticks, float inv, bool* saturated) {
  const float tx = (float)ticks[0], ty = (float)ticks[1], tz = (float)ticks[2];
  const float m = fmaxf(fmaxf(fabsf(tx), fabsf(ty)), fabsf(tz));
  *saturated = m >= 1073741824.0f;
  return make_float3(tx * inv, ty * inv, tz * inv);
}

}  // namespace pf
