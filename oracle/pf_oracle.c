/*
 * pf_oracle.c — CPU oracle for the MI355X particle solver.
 *
 * TEST INFRASTRUCTURE ONLY. Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load this library, and only as the checker
 * or the timed CPU baseline, never as a product path.
 *
 * It restates the reference's per-substep solver on the CPU:
 *   - Morton bucket key and stable counting sort
 *     (src/core/hashGrid/mortonHash.ts:39-79, cellIndex.ts:113-170, sort.ts:55-195);
 *   - the 27-bucket neighbour walk (query.ts:24-70) and the 64-neighbour list
 *     (neighborList.ts:45-72);
 *   - the full SimLoop substep for fluids, primitive colliders and global
 *     shape matching (orc_sim.h), in FP32 (operand order of the GPU kernels)
 *     and in FP64.
 * The reference itself is TypeScript/WGSL that cannot run here (no tsc,
 * three.js or WebGPU), so nothing under /root/reference is compiled.
 */
#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

/* ---------------------------------------------------------------------- */
/* integer part (mortonHash.ts, cellIndex.ts)                              */
/* ---------------------------------------------------------------------- */
uint32_t orc_part1by2(uint32_t n) {
  n &= 0x000003ffu;
  n = (n | (n << 16)) & 0xff0000ffu;
  n = (n | (n << 8)) & 0x0300f00fu;
  n = (n | (n << 4)) & 0x030c30c3u;
  n = (n | (n << 2)) & 0x09249249u;
  return n;
}

uint32_t orc_morton(int32_t cx, int32_t cy, int32_t cz) {
  uint32_t ux = (uint32_t)cx + 512u, uy = (uint32_t)cy + 512u, uz = (uint32_t)cz + 512u;
  return orc_part1by2(ux) | (orc_part1by2(uy) << 1) | (orc_part1by2(uz) << 2);
}

/* WGSL f32 -> i32 conversion saturates; NaN is taken as 0. */
int32_t orc_f2i_sat(double f) {
  if (!(f == f)) return 0;
  if (f >= 2147483648.0) return 2147483647;
  if (f < -2147483648.0) return (int32_t)(-2147483647 - 1);
  return (int32_t)f;
}

/* ---------------------------------------------------------------------- */
/* half floats (SDFCollider.ts:97-101)                                     */
/* ---------------------------------------------------------------------- */
/* three@0.184.0 DataUtils.toHalfFloat (not vendored in the reference): the
 * value is clamped to +-65504, then converted with the base/shift tables of
 * van der Zijp's "Fast Half Float Conversions" (2008), which truncate the
 * mantissa: half = base[sign|exponent] + (mantissa >> shift[sign|exponent]). */
static uint16_t orc_half_base[512];
static uint8_t orc_half_shift[512];
static int orc_half_ready = 0;

static void orc_half_tables(void) {
  for (int i = 0; i < 256; ++i) {
    int e = i - 127;
    uint16_t b;
    uint8_t sh;
    if (e < -27) { b = 0x0000; sh = 24; }
    else if (e < -14) { b = (uint16_t)(0x0400 >> (-e - 14)); sh = (uint8_t)(-e - 1); }
    else if (e <= 15) { b = (uint16_t)((e + 15) << 10); sh = 13; }
    else if (e < 128) { b = 0x7c00; sh = 24; }
    else { b = 0x7c00; sh = 13; }
    orc_half_base[i] = b;
    orc_half_base[i | 0x100] = (uint16_t)(b | 0x8000);
    orc_half_shift[i] = sh;
    orc_half_shift[i | 0x100] = sh;
  }
  orc_half_ready = 1;
}

uint16_t orc_float_to_half(float v) {
  if (!orc_half_ready) orc_half_tables();
  if (v > 65504.0f) v = 65504.0f;
  if (v < -65504.0f) v = -65504.0f;
  uint32_t f;
  memcpy(&f, &v, 4);
  uint32_t e = (f >> 23) & 0x1ffu;
  return (uint16_t)(orc_half_base[e] + ((f & 0x007fffffu) >> orc_half_shift[e]));
}

/* exact value of a half float */
double orc_half_to_double(uint16_t h) {
  int sgn = h >> 15, e = (h >> 10) & 0x1f, m = h & 0x3ff;
  double v;
  if (e == 0) v = ldexp((double)m, -24);
  else if (e == 31) v = m ? NAN : INFINITY;
  else v = ldexp((double)(m | 0x400), e - 25);
  return sgn ? -v : v;
}

/* Bucket of one FP32 position: (p - o) / cell in FP32, floor, saturate,
 * clamp to +-2^30, Morton, mask (cellIndex.ts:113-170). */
uint32_t orc_bucket(const float p[3], const float o[3], float cell, uint32_t H, int* over) {
  int32_t c[3];
  for (int a = 0; a < 3; ++a) {
    float rel = (p[a] - o[a]) / cell;
    int32_t raw = orc_f2i_sat((double)floorf(rel));
    int32_t cc = raw < -(1 << 30) ? -(1 << 30) : (raw > (1 << 30) - 1 ? (1 << 30) - 1 : raw);
    if (raw != cc || cc < -512 || cc >= 512) *over = 1;
    c[a] = cc;
  }
  return orc_morton(c[0], c[1], c[2]) & (H - 1);
}

/* HashGrid rebuild (HashGrid.ts:195-204) as a stable counting sort.
 * counts/cell_start/cell_end hold Hpad = ceil(H/1024)*1024 entries. */
int orc_grid_build(const float* pos4, uint32_t n, float cell, const float origin[3], uint32_t H,
                   uint32_t* cell_index, uint32_t* counts, uint32_t* cell_start, uint32_t* cell_end,
                   uint32_t* sorted_indices, uint32_t* sorted_keys) {
  int over = 0;
  uint32_t Hpad = (uint32_t)(((uint64_t)H + 1023) / 1024 * 1024);
  uint32_t* cnt = (uint32_t*)calloc(Hpad, sizeof(uint32_t));
  uint32_t* key = cell_index ? cell_index : (uint32_t*)malloc(sizeof(uint32_t) * (n ? n : 1));
  for (uint32_t i = 0; i < n; ++i) {
    key[i] = orc_bucket(pos4 + 4 * (size_t)i, origin, cell, H, &over);
    cnt[key[i]]++;
  }
  uint32_t* st = (uint32_t*)malloc(sizeof(uint32_t) * ((size_t)Hpad + 1));
  uint32_t run = 0;
  for (uint32_t b = 0; b < Hpad; ++b) {
    st[b] = run;
    run += cnt[b];
  }
  st[Hpad] = run;
  if (counts) memcpy(counts, cnt, sizeof(uint32_t) * Hpad);
  if (cell_start) memcpy(cell_start, st, sizeof(uint32_t) * Hpad);
  if (cell_end)
    for (uint32_t b = 0; b < Hpad; ++b) cell_end[b] = st[b] + cnt[b];
  for (uint32_t i = 0; i < n; ++i) {
    uint32_t slot = st[key[i]]++;
    if (sorted_indices) sorted_indices[slot] = i;
    if (sorted_keys) sorted_keys[slot] = key[i];
  }
  free(st);
  free(cnt);
  if (!cell_index) free(key);
  return over;
}

/* emitForEachNeighbor (query.ts:24-70): for each query point, candidates
 * visited in the 27 buckets (own index excluded) and those within r. */
void orc_grid_query_counts(const float* pos4, uint32_t n, float cell, const float origin[3], uint32_t H,
                           const float* query4, uint32_t nq, float radius, uint32_t* cand,
                           uint32_t* within) {
  uint32_t Hpad = (uint32_t)(((uint64_t)H + 1023) / 1024 * 1024);
  uint32_t* cs = (uint32_t*)malloc(sizeof(uint32_t) * Hpad);
  uint32_t* ce = (uint32_t*)malloc(sizeof(uint32_t) * Hpad);
  uint32_t* perm = (uint32_t*)malloc(sizeof(uint32_t) * (n ? n : 1));
  orc_grid_build(pos4, n, cell, origin, H, NULL, NULL, cs, ce, perm, NULL);
  float r2 = radius * radius;
  for (uint32_t i = 0; i < nq; ++i) {
    const float* x = query4 + 4 * (size_t)i;
    int32_t q[3];
    for (int a = 0; a < 3; ++a) q[a] = orc_f2i_sat((double)floorf((x[a] - origin[a]) / cell));
    uint32_t visited[27], nv = 0, nc = 0, nw = 0;
    for (int dz = -1; dz <= 1; ++dz)
      for (int dy = -1; dy <= 1; ++dy)
        for (int dx = -1; dx <= 1; ++dx) {
          uint32_t b = orc_morton((int32_t)((uint32_t)q[0] + (uint32_t)dx), (int32_t)((uint32_t)q[1] + (uint32_t)dy),
                                  (int32_t)((uint32_t)q[2] + (uint32_t)dz)) & (H - 1);
          int seen = 0;
          for (uint32_t k = 0; k < nv; ++k) seen |= visited[k] == b;
          visited[nv++] = b;
          if (seen) continue;
          for (uint32_t t = cs[b]; t < ce[b]; ++t) {
            uint32_t j = perm[t];
            if (j == i) continue;
            ++nc;
            const float* y = pos4 + 4 * (size_t)j;
            float ddx = x[0] - y[0], ddy = x[1] - y[1], ddz = x[2] - y[2];
            if ((ddx * ddx + ddy * ddy) + ddz * ddz < r2) ++nw;
          }
        }
    if (cand) cand[i] = nc;
    if (within) within[i] = nw;
  }
  free(cs);
  free(ce);
  free(perm);
}

/* Atomic wrap-around add into a fixed-point accumulator word: a scatter
 * from parallel particle loops (integer sums do not depend on order). */
static inline void orc_acc_add(int32_t* a, int32_t v) {
  uint32_t* u = (uint32_t*)a;
#pragma omp atomic
  *u += (uint32_t)v;
}

/* Threads the oracle's particle loops use (0: OpenMP's default). */
void orc_set_threads(int n) { omp_set_num_threads(n > 0 ? n : omp_get_num_procs()); }
int orc_get_threads(void) { return omp_get_max_threads(); }

/* ---------------------------------------------------------------------- */
/* FP32 mirror and FP64 restatement of the substep                        */
/* ---------------------------------------------------------------------- */
#define REAL float
#define SFX(name) name##_f32
#define R(x) ((float)(x))
#define RSQRT sqrtf
#define RFLOOR floorf
#define RFMAX fmaxf
#define RFMIN fminf
#define RFABS fabsf
#define RSIN sinf
#define RCOS cosf
#define RATAN2 atan2f
#define RRINT rintf
#define REXP expf
#include "orc_sim.h"
#undef REAL
#undef SFX
#undef R
#undef RATAN2
#undef RSQRT
#undef RFLOOR
#undef RFMAX
#undef RFMIN
#undef RFABS
#undef RSIN
#undef RCOS
#undef RRINT
#undef REXP

#define REAL double
#define SFX(name) name##_f64
#define R(x) ((double)(x))
#define RSQRT sqrt
#define RFLOOR floor
#define RFMAX fmax
#define RFMIN fmin
#define RFABS fabs
#define RSIN sin
#define RCOS cos
#define RATAN2 atan2
#define RRINT rint
#define REXP exp
#include "orc_sim.h"
#undef REAL
#undef SFX
#undef R
#undef RATAN2

/* ---------------------------------------------------------------------- */
/* exported scene API (same entry points for both precisions)              */
/* ---------------------------------------------------------------------- */
#define ORC_SCENE_API(SUF, T)                                                                      \
  void* orc_scene_create_##SUF(uint32_t n, double radius) {                                        \
    scene_##SUF* s = (scene_##SUF*)calloc(1, sizeof(scene_##SUF));                                 \
    s->n = n;                                                                                      \
    s->radius = (T)radius;                                                                         \
    s->x = (T*)calloc(3 * (size_t)n, sizeof(T));                                                   \
    s->xs = (T*)calloc(3 * (size_t)n, sizeof(T));                                                  \
    s->v = (T*)calloc(3 * (size_t)n, sizeof(T));                                                   \
    s->w = (T*)calloc(n, sizeof(T));                                                               \
    s->rho = (T*)calloc(n, sizeof(T));                                                             \
    s->bvol = (T*)calloc(n, sizeof(T));                                                            \
    s->key = (uint32_t*)calloc(n, sizeof(uint32_t));                                               \
    s->perm = (uint32_t*)calloc(n, sizeof(uint32_t));                                              \
    s->plam = (T*)calloc(64 * (size_t)n, sizeof(T));                                               \
    s->pacc = (int32_t*)calloc(3 * (size_t)n, sizeof(int32_t));                                    \
    s->vacc = (int32_t*)calloc(3 * (size_t)n, sizeof(int32_t));                                    \
    s->iacc = (int32_t*)calloc(3 * (size_t)n, sizeof(int32_t));                                    \
    s->racc = (int32_t*)calloc(3 * (size_t)n, sizeof(int32_t));                                    \
    s->rest = (T*)calloc(3 * (size_t)n, sizeof(T));                                                \
    s->weights = (T*)calloc(n, sizeof(T));                                                         \
    s->slam = (T*)calloc(3 * (size_t)n, sizeof(T));                                                \
    s->radius_d = radius;                                                                          \
    s->grp = (uint32_t*)calloc(n, sizeof(uint32_t));                                               \
    s->xs_grid = (T*)calloc(3 * (size_t)n, sizeof(T));                                             \
    s->q = (T*)calloc(4 * (size_t)n, sizeof(T));                                                   \
    s->qp = (T*)calloc(4 * (size_t)n, sizeof(T));                                                  \
    s->om = (T*)calloc(4 * (size_t)n, sizeof(T));                                                  \
    s->lcomp = (T*)calloc(n, sizeof(T));                                                           \
    for (size_t k = 0; k < n; ++k) { s->q[4 * k + 3] = (T)1; s->qp[4 * k + 3] = (T)1; }            \
    s->substeps = 4;                                                                               \
    s->iterations = 2;                                                                             \
    s->g[1] = (T)-9.81;                                                                            \
    s->pos_scale = (T)floor(1073741824.0 / 10.0);                                                  \
    s->pos_inv = (T)(1.0 / floor(1073741824.0 / 10.0));                                            \
    s->vel_scale = (T)floor(1073741824.0 / 50.0);                                                  \
    s->vel_inv = (T)(1.0 / floor(1073741824.0 / 50.0));                                            \
    return s;                                                                                      \
  }                                                                                                \
  void orc_scene_destroy_##SUF(void* p) {                                                          \
    scene_##SUF* s = (scene_##SUF*)p;                                                              \
    if (!s) return;                                                                                \
    free(s->x); free(s->xs); free(s->v); free(s->w); free(s->rho); free(s->bvol);                  \
    free(s->key); free(s->perm); free(s->start); free(s->plam); free(s->pacc); free(s->vacc);      \
    free(s->iacc); free(s->racc); free(s->rest); free(s->weights); free(s->slam); free(s->sdflam); \
    for (uint32_t k = 0; k < s->nsdf; ++k) free(s->sdf[k].grid);                                   \
    free(s->q); free(s->qp); free(s->om); free(s->lcomp); free(s->ledges); free(s->csr_off);      \
    free(s->csr_idx); free(s->rcen); free(s->gcen); free(s->grot); free(s->plam4); free(s->lacc); \
    free(s->grp); free(s->xs_grid); free(s->ct_i); free(s->ct_j); free(s->ct_ln); free(s->ct_lt);  \
    free(s->ct_n);                                                                                 \
    for (uint32_t f = 0; f < s->nfluids; ++f) {                                                    \
      fluid_##SUF* F = &s->fluid[f];                                                               \
      free(F->nl); free(F->nlc); free(F->normal); free(F->lambda); free(F->dx);                    \
      free(F->omega); free(F->omlen); free(F->eta); free(F->dv);                                   \
      free(F->isv_rhs); free(F->isv_a); free(F->isv_b);                                            \
    }                                                                                              \
    free(s);                                                                                       \
  }                                                                                                \
  /* positions/velocities: xyz triples (float); inv_mass per particle */                          \
  void orc_scene_set_particles_##SUF(void* p, const float* x3, const float* v3, const float* w) {   \
    scene_##SUF* s = (scene_##SUF*)p;                                                              \
    for (size_t k = 0; k < 3 * (size_t)s->n; ++k) {                                                \
      s->x[k] = (T)x3[k];                                                                          \
      s->xs[k] = (T)x3[k];                                                                         \
      s->v[k] = v3 ? (T)v3[k] : (T)0;                                                              \
    }                                                                                              \
    for (uint32_t i = 0; i < s->n; ++i) s->w[i] = w ? (T)w[i] : (T)1;                              \
  }                                                                                                \
  void orc_scene_set_loop_##SUF(void* p, uint32_t substeps, uint32_t iterations, const double g[3], \
                                double cell, uint32_t H, const double origin[3]) {                 \
    scene_##SUF* s = (scene_##SUF*)p;                                                              \
    s->substeps = substeps;                                                                        \
    s->iterations = iterations;                                                                    \
    for (int a = 0; a < 3; ++a) {                                                                  \
      s->g[a] = (T)g[a];                                                                           \
      s->origin[a] = origin ? (T)origin[a] : (T)0;                                                 \
    }                                                                                              \
    s->cell = (T)cell;                                                                             \
    s->H = H;                                                                                      \
    free(s->start);                                                                                \
    s->start = (uint32_t*)calloc((size_t)H + 1, sizeof(uint32_t));                                 \
  }                                                                                                \
  /* FluidSystem.ts:90-127 + build 200-309; sets inv mass and density of the range */             \
  int orc_scene_add_fluid_##SUF(void* p, uint32_t start, uint32_t count, double rho0,              \
                                double spacing, double h, double compliance, uint32_t effects,     \
                                double visc, double vort, double st, double adh) {                 \
    scene_##SUF* s = (scene_##SUF*)p;                                                              \
    if (s->nfluids >= 4) return -1;                                                                \
    fluid_##SUF* F = &s->fluid[s->nfluids++];                                                      \
    double mass = rho0 * spacing * spacing * spacing;                                              \
    F->start = start; F->count = count; F->effects = effects;                                      \
    F->rho0 = (T)rho0; F->pv = (T)(spacing * spacing * spacing); F->mass = (T)mass; F->inv_mass = (T)1 / F->mass;                \
    F->h = (T)h; F->h2 = (T)(h * h); F->h_d = h;                                                   \
    F->poly6 = (T)(315.0 / (64.0 * M_PI * pow(h, 9)));                                              \
    F->spiky = (T)(45.0 / (M_PI * pow(h, 6)));                                                      \
    F->compliance = (T)compliance;                                                                 \
    F->gamma = (T)st; F->visc = (T)visc; F->vort = (T)vort; F->adhesion = (T)adh;                  \
    F->coh_coef = (T)(32.0 / (M_PI * pow(h, 9)));                                                   \
    F->coh_h6_64 = (T)(pow(h, 6) / 64.0);                                                           \
    F->adh_coef = (T)(0.007 / pow(h, 3.25));                                                        \
    F->adh_m4_over_h = (T)(-4.0 / h);                                                               \
    F->adh_2h = (T)(2.0 * h);                                                                       \
    F->imp_scale = (T)floor(1073741824.0 / 50.0);                                                  \
    F->imp_inv = (T)(1.0 / floor(1073741824.0 / 50.0));                                           \
    F->rea_scale = (T)floor(1073741824.0 / 10.0);                                                  \
    F->rea_inv = (T)(1.0 / floor(1073741824.0 / 10.0));                                            \
    size_t n = s->n;                                                                               \
    F->nl = (uint32_t*)calloc(64 * n, sizeof(uint32_t));                                           \
    F->nlc = (uint32_t*)calloc(n, sizeof(uint32_t));                                               \
    F->normal = (T*)calloc(3 * n, sizeof(T));                                                      \
    F->lambda = (T*)calloc(n, sizeof(T));                                                          \
    F->dx = (T*)calloc(3 * n, sizeof(T));                                                          \
    F->omega = (T*)calloc(3 * n, sizeof(T));                                                       \
    F->omlen = (T*)calloc(n, sizeof(T));                                                           \
    F->eta = (T*)calloc(3 * n, sizeof(T));                                                         \
    F->dv = (T*)calloc(3 * n, sizeof(T));                                                          \
    /* invMass = 1 / mass stored as f32 (FluidSystem.ts:118) */                                   \
    float im = (float)(1.0 / mass);                                                                \
    for (uint32_t i = start; i < start + count; ++i) {                                             \
      s->w[i] = (T)im;                                                                             \
      s->rho[i] = (T)(float)rho0;                                                                  \
    }                                                                                              \
    return 0;                                                                                      \
  }                                                                                                \
  /* FluidSystem.addBoundary (FluidSystem.ts:171-183): range and dynamic flag of fluid f */        \
  int orc_scene_add_boundary_##SUF(void* p, uint32_t f, uint32_t start, uint32_t count, int dyn) { \
    scene_##SUF* s = (scene_##SUF*)p;                                                              \
    if (f >= s->nfluids || s->fluid[f].nb >= 8) return -1;                                         \
    fluid_##SUF* F = &s->fluid[f];                                                                 \
    F->bstart[F->nb] = start; F->bcount[F->nb] = count; F->bdyn[F->nb] = dyn ? 1u : 0u;            \
    F->nb++;                                                                                       \
    return 0;                                                                                      \
  }                                                                                                \
  /* SDFCollider (SDFCollider.ts:68-311): voxels as float (converted here), placement   \
   * and friction as the f32 uniforms the reference uploads. rot: row-major 3x3. */     \
  int orc_scene_add_sdf_##SUF(void* p, const float* data, const uint32_t res[3], const double origin[3], \
                              const double voxel[3], const double pos[3], const double rot9[9],        \
                              double scale, double thickness, double mus, double muk) {                \
    scene_##SUF* s = (scene_##SUF*)p;                                                              \
    if (s->nsdf >= 4) return -1;                                                                   \
    sdfc_##SUF* D = &s->sdf[s->nsdf++];                                                             \
    size_t nv = (size_t)res[0] * res[1] * res[2];                                                  \
    D->grid = (uint16_t*)malloc(nv * sizeof(uint16_t));                                            \
    for (size_t k = 0; k < nv; ++k) D->grid[k] = orc_float_to_half(data[k]);                        \
    D->nx = (int)res[0]; D->ny = (int)res[1]; D->nz = (int)res[2];                                 \
    for (int a = 0; a < 3; ++a) {                                                                  \
      D->res[a] = (T)(float)res[a];                                                                \
      D->inv_res[a] = (T)(float)(1.0 / (double)res[a]);                                            \
      D->origin[a] = (T)(float)origin[a];                                                          \
      D->voxel[a] = (T)(float)voxel[a];                                                            \
      D->inv_voxel[a] = (T)(float)(1.0 / voxel[a]);                                                \
      D->pos[a] = (T)(float)pos[a];                                                                \
    }                                                                                              \
    for (int r = 0; r < 3; ++r)                                                                    \
      for (int c = 0; c < 3; ++c) {                                                                \
        D->rot[3 * r + c] = (T)(float)rot9[3 * r + c];                                             \
        D->inv_rot[3 * r + c] = (T)(float)rot9[3 * c + r];                                         \
      }                                                                                            \
    D->scale = (T)(float)scale; D->inv_scale = (T)(float)(1.0 / scale);                            \
    D->thickness = (T)(float)thickness; D->mus = (T)(float)mus; D->muk = (T)(float)muk;            \
    if (!s->sdflam) s->sdflam = (T*)calloc(2 * 4 * (size_t)s->n, sizeof(T));                       \
    return 0;                                                                                      \
  }                                                                                                \
  /* phi and gradient of SDF collider k at a world point (sdf.ts:107-146) */                       \
  void orc_scene_sdf_sample_##SUF(void* p, uint32_t k, const double x[3], double* phi, double g[3]) { \
    scene_##SUF* s = (scene_##SUF*)p;                                                              \
    T xx[3] = {(T)(float)x[0], (T)(float)x[1], (T)(float)x[2]}, ph, gg[3];                         \
    sdf_sample_##SUF(&s->sdf[k], xx, &ph, gg);                                                     \
    *phi = (double)ph;                                                                             \
    for (int a = 0; a < 3; ++a) g[a] = (double)gg[a];                                              \
  }                                                                                                \
  /* boundary volumes, in particle order */                                                        \
  void orc_scene_get_bvol_##SUF(void* p, double* out) {                                            \
    scene_##SUF* s = (scene_##SUF*)p;                                                              \
    for (uint32_t i = 0; i < s->n; ++i) out[i] = (double)s->bvol[i];                               \
  }                                                                                                \
  /* PrimitiveSet add* after the host-side double setup; data as uploaded (f32 values) */         \
  int orc_scene_add_prim_##SUF(void* p, uint32_t kind, uint32_t flags, const double d0[4],         \
                               const double d1[4], const double vel[4], const double rot[4],       \
                               const double spin[3]) {                                             \
    scene_##SUF* s = (scene_##SUF*)p;                                                              \
    if (s->nprim >= 32) return -1;                                                                 \
    prim_##SUF* P = &s->prim[s->nprim++];                                                          \
    P->kind = kind; P->flags = flags;                                                              \
    for (int a = 0; a < 4; ++a) {                                                                  \
      P->d0[a] = (T)(float)d0[a]; P->d1[a] = (T)(float)d1[a];                                      \
      P->vel[a] = (T)(float)vel[a]; P->rot[a] = (T)(float)rot[a];                                  \
    }                                                                                              \
    for (int a = 0; a < 3; ++a) P->spin[a] = (T)(float)spin[a];                                    \
    return 0;                                                                                      \
  }                                                                                                \
  /* ViscositySolver listed after fluid f (ViscositySolver.ts:49-58, 112) */                      \
  int orc_scene_add_viscosity_##SUF(void* p, uint32_t f, double nu, uint32_t iterations) {         \
    scene_##SUF* s = (scene_##SUF*)p;                                                              \
    if (f >= s->nfluids || iterations < 1 || iterations > 64) return -1;                          \
    fluid_##SUF* F = &s->fluid[f];                                                                 \
    F->isv_iters = iterations;                                                                     \
    F->isv_nu = (T)nu;                                                                             \
    F->isv_c10 = (T)(10.0 / (F->h_d * F->h_d));                                                    \
    F->isv_rhs = (T*)calloc(3 * (size_t)s->n, sizeof(T));                                          \
    F->isv_a = (T*)calloc(3 * (size_t)s->n, sizeof(T));                                            \
    F->isv_b = (T*)calloc(3 * (size_t)s->n, sizeof(T));                                            \
    return 0;                                                                                      \
  }                                                                                                \
  /* replace primitive k (a collider moved between steps) */                                      \
  int orc_scene_set_prim_##SUF(void* p, uint32_t k, uint32_t kind, uint32_t flags, const double d0[4], \
                               const double d1[4], const double vel[4], const double rot[4],       \
                               const double spin[3]) {                                             \
    scene_##SUF* s = (scene_##SUF*)p;                                                              \
    if (k >= s->nprim) return -1;                                                                  \
    prim_##SUF* P = &s->prim[k];                                                                   \
    P->kind = kind; P->flags = flags;                                                              \
    for (int a = 0; a < 4; ++a) {                                                                  \
      P->d0[a] = (T)(float)d0[a]; P->d1[a] = (T)(float)d1[a];                                      \
      P->vel[a] = (T)(float)vel[a]; P->rot[a] = (T)(float)rot[a];                                  \
    }                                                                                              \
    for (int a = 0; a < 3; ++a) P->spin[a] = (T)(float)spin[a];                                    \
    return 0;                                                                                      \
  }                                                                                                \
  /* one global shape-matching body; rest offsets and weights as SoftbodySystem computes them */  \
  int orc_scene_add_body_##SUF(void* p, uint32_t start, uint32_t count, double compliance,        \
                               const float* rest3, const float* weights) {                         \
    scene_##SUF* s = (scene_##SUF*)p;                                                              \
    if (s->nbodies >= 4096) return -1;                                                             \
    uint32_t b = s->nbodies++;                                                                     \
    s->bstart[b] = start; s->bcount[b] = count; s->bcomp[b] = (T)(float)compliance;                \
    for (uint32_t i = 0; i < count; ++i) s->lcomp[start + i] = (T)(float)compliance;               \
    for (uint32_t i = 0; i < count; ++i) {                                                         \
      for (int a = 0; a < 3; ++a) s->rest[3 * (size_t)(start + i) + a] = (T)rest3[3 * (size_t)i + a]; \
      s->weights[start + i] = (T)weights[i];                                                       \
    }                                                                                              \
    return 0;                                                                                      \
  }                                                                                                \
  /* edges of the last added body, local pairs (SoftbodyDef.edges); switches the scene to local   \
   * shape matching (SoftbodySystem shapeMatching: 'local') */                                     \
  int orc_scene_add_body_edges_##SUF(void* p, const uint32_t* edges, uint32_t nedges) {            \
    scene_##SUF* s = (scene_##SUF*)p;                                                              \
    if (!s->nbodies) return -1;                                                                    \
    uint32_t b = s->nbodies - 1, st = s->bstart[b];                                                \
    if (s->nledges + nedges > s->capledges) {                                                      \
      s->capledges = 2 * (s->nledges + nedges);                                                    \
      s->ledges = (uint32_t*)realloc(s->ledges, 2 * (size_t)s->capledges * sizeof(uint32_t));      \
    }                                                                                              \
    for (uint32_t e = 0; e < nedges; ++e) {                                                        \
      s->ledges[2 * (size_t)(s->nledges + e)] = st + edges[2 * e];                                 \
      s->ledges[2 * (size_t)(s->nledges + e) + 1] = st + edges[2 * e + 1];                         \
    }                                                                                              \
    s->nledges += nedges;                                                                          \
    s->local = 1;                                                                                  \
    return 0;                                                                                      \
  }                                                                                                \
  /* orientation, predicted orientation, angular velocity (xyzw per particle) */                  \
  void orc_scene_get_orientation_##SUF(void* p, double* q4, double* qp4, double* om4) {           \
    scene_##SUF* s = (scene_##SUF*)p;                                                              \
    for (size_t k = 0; k < 4 * (size_t)s->n; ++k) {                                                \
      if (q4) q4[k] = (double)s->q[k];                                                             \
      if (qp4) qp4[k] = (double)s->qp[k];                                                          \
      if (om4) om4[k] = (double)s->om[k];                                                          \
    }                                                                                              \
  }                                                                                                \
  void orc_scene_set_groups_##SUF(void* p, const uint32_t* g) {                                    \
    scene_##SUF* s = (scene_##SUF*)p;                                                              \
    memcpy(s->grp, g, sizeof(uint32_t) * s->n);                                                    \
  }                                                                                                \
  /* SimLoop contact option (loop.ts:23-31, 346-418): contact distance 2r,                       \
   * candidate radius 2r * 1.1 (generate.ts:19), quiet: [start, count] pairs */                   \
  int orc_scene_set_contact_##SUF(void* p, double mus, double muk, uint32_t max_contacts,         \
                                  const uint32_t* quiet, uint32_t nquiet) {                        \
    scene_##SUF* s = (scene_##SUF*)p;                                                              \
    if (nquiet > 8 || max_contacts == 0) return -1;                                                \
    double cr = 2.0 * s->radius_d * 1.1;                                                           \
    s->contact = 1;                                                                                \
    s->ct_dist = (T)(2.0 * s->radius_d);                                                           \
    s->ct_cr2 = (T)(cr * cr);                                                                      \
    s->ct_mus = (T)mus;                                                                            \
    s->ct_muk = (T)muk;                                                                            \
    s->ct_max = max_contacts;                                                                      \
    s->nquiet = nquiet;                                                                            \
    for (uint32_t q = 0; q < nquiet; ++q) { s->qstart[q] = quiet[2 * q]; s->qcount[q] = quiet[2 * q + 1]; } \
    s->ct_i = (uint32_t*)calloc(max_contacts, sizeof(uint32_t));                                   \
    s->ct_j = (uint32_t*)calloc(max_contacts, sizeof(uint32_t));                                   \
    s->ct_ln = (int32_t*)calloc(max_contacts, sizeof(int32_t));                                    \
    s->ct_lt = (int32_t*)calloc(max_contacts, sizeof(int32_t));                                    \
    s->ct_n = (T*)calloc(3 * (size_t)max_contacts, sizeof(T));                                     \
    return 0;                                                                                      \
  }                                                                                                \
  /* readbackCount / readbackPairs (ContactBuffer.ts:74-97): returns the emitted                 \
   * count; pairs receives min(count, maxContacts, cap) (i, j) pairs */                           \
  uint32_t orc_scene_contacts_##SUF(void* p, uint32_t* pairs, uint32_t cap) {                      \
    scene_##SUF* s = (scene_##SUF*)p;                                                              \
    uint32_t used = s->ct_count < s->ct_max ? s->ct_count : s->ct_max;                             \
    for (uint32_t c = 0; c < used && c < cap; ++c) { pairs[2 * c] = s->ct_i[c]; pairs[2 * c + 1] = s->ct_j[c]; } \
    return s->ct_count;                                                                            \
  }                                                                                                \
  /* new SurfaceField({min, max, voxelBudget, motionStretch, solids, colliders})                 \
   * (SurfaceField.ts:116-244): the constructor's double arithmetic, then one                     \
   * frame of its kernels; out: 4 values per voxel, x fastest; res, voxel: the                     \
   * grid. radius / men_h / men_w: the uniforms (r 1.25, r 1.6, r 2.5 by default). */            \
  int orc_scene_surface_field_##SUF(void* p, const double mn[3], const double mx[3], double budget,   \
                                    double stretch, double spacing, const uint32_t ranges[4],       \
                                    int wet, const double uni[3], T* out, uint32_t res[3],          \
                                    double* voxel_out) {                                            \
    scene_##SUF* s = (scene_##SUF*)p;                                                              \
    double sz[3] = {mx[0] - mn[0], mx[1] - mn[1], mx[2] - mn[2]};                                 \
    if (!(sz[0] > 0 && sz[1] > 0 && sz[2] > 0)) return -1;                                         \
    double voxel = fmax(s->radius_d, cbrt((sz[0] * sz[1] * sz[2]) / budget));                       \
    fieldp_##SUF F;                                                                                \
    memset(&F, 0, sizeof F);                                                                       \
    for (int a = 0; a < 3; ++a) {                                                                  \
      F.n[a] = (uint32_t)ceil(sz[a] / voxel);                                                      \
      F.origin[a] = (T)mn[a];                                                                      \
      F.dims[a] = (T)F.n[a];                                                                       \
      res[a] = F.n[a];                                                                             \
    }                                                                                              \
    *voxel_out = voxel;                                                                            \
    F.count = F.n[0] * F.n[1] * F.n[2];                                                            \
    if (!out) return 0;                                                                            \
    F.voxel = (T)voxel;                                                                            \
    F.inv_voxel = (T)(1.0 / voxel);                                                                \
    double sigma = (0.75 * spacing) / voxel;                                                       \
    int taps = (int)fmin(6.0, fmax(2.0, ceil(2.5 * sigma)));                                       \
    double wsum = 0, wt[13];                                                                       \
    for (int k = 0; k <= 2 * taps; ++k) {                                                          \
      wt[k] = exp(-((double)(k - taps) * (double)(k - taps)) / (2 * sigma * sigma));              \
      wsum += wt[k];                                                                               \
    }                                                                                              \
    for (int k = 0; k <= 2 * taps; ++k) F.weights[k] = (T)(wt[k] / wsum);                         \
    F.taps = (uint32_t)taps;                                                                       \
    F.samples = stretch > 0 ? 3u : 1u;                                                             \
    F.stretch = (T)stretch;                                                                        \
    F.reach3 = (T)(3 * spacing);                                                                   \
    F.inv_samples = (T)(1.0 / F.samples);                                                          \
    double vs = (spacing / voxel) * (spacing / voxel) * (spacing / voxel);                         \
    F.vol_scale = (T)vs;                                                                           \
    F.solid_scale = (T)(vs * 1.5);                                                                 \
    F.fluid_start = ranges[0]; F.fluid_count = ranges[1];                                          \
    F.solid_start = ranges[2]; F.solid_count = ranges[3];                                          \
    F.wet = wet ? 1u : 0u;                                                                         \
    F.radius = (T)uni[0]; F.men_h = (T)uni[1]; F.men_w = (T)uni[2];                                \
    surface_field_##SUF(s, &F, out);                                                               \
    return 0;                                                                                      \
  }                                                                                                \
  void orc_scene_step_##SUF(void* p, double dt) { step_##SUF((scene_##SUF*)p, dt); }               \
  /* state out as double: x, x*, v as xyz triples, density */                                     \
  void orc_scene_get_##SUF(void* p, double* x3, double* xs3, double* v3, double* rho) {           \
    scene_##SUF* s = (scene_##SUF*)p;                                                              \
    for (size_t k = 0; k < 3 * (size_t)s->n; ++k) {                                                \
      if (x3) x3[k] = (double)s->x[k];                                                             \
      if (xs3) xs3[k] = (double)s->xs[k];                                                          \
      if (v3) v3[k] = (double)s->v[k];                                                             \
    }                                                                                              \
    if (rho) for (uint32_t i = 0; i < s->n; ++i) rho[i] = (double)s->rho[i];                       \
  }                                                                                                \
  void orc_scene_flags_##SUF(void* p, int* out5) {                                                 \
    scene_##SUF* s = (scene_##SUF*)p;                                                              \
    out5[0] = s->flag_grid; out5[1] = s->flag_nl; out5[2] = s->flag_pos;                           \
    out5[3] = s->flag_vel; out5[4] = s->flag_imp;                                                  \
  }                                                                                                \
  void orc_scene_body_frame_##SUF(void* p, uint32_t b, double c[3], double r9[9]) {               \
    scene_##SUF* s = (scene_##SUF*)p;                                                              \
    for (int a = 0; a < 3; ++a) c[a] = (double)s->centers[b][a];                                   \
    for (int k = 0; k < 9; ++k) r9[k] = (double)s->rot[b][k];                                      \
  }                                                                                                \
  void orc_polar_##SUF(const double a9[9], double r9[9]) {                                         \
    T a[9], r[9];                                                                                  \
    for (int k = 0; k < 9; ++k) a[k] = (T)a9[k];                                                   \
    polar_##SUF(a, r);                                                                             \
    for (int k = 0; k < 9; ++k) r9[k] = (double)r[k];                                              \
  }                                                                                                \
  /* neighbour lists of the last substep, in walk order (row-major, 64 per particle) */           \
  void orc_scene_neighbors_##SUF(void* p, uint32_t f, uint32_t* nl, uint32_t* cnt) {              \
    scene_##SUF* s = (scene_##SUF*)p;                                                              \
    fluid_##SUF* F = &s->fluid[f];                                                                 \
    memcpy(nl, F->nl, sizeof(uint32_t) * 64 * (size_t)s->n);                                       \
    memcpy(cnt, F->nlc, sizeof(uint32_t) * (size_t)s->n);                                          \
  }                                                                                                \
  /* fluid f's per-pass fields as the last substep left them, by particle id:                     \
   * lambda and position delta of the last iteration (lambda.ts, positionDelta.ts),               \
   * colour normal (cohesion.ts:25-39), omega, |omega|, eta and the XSPH change                   \
   * (velocity.ts:36-92); any output may be null */                                              \
  void orc_scene_fluid_fields_##SUF(void* p, uint32_t f, double* lam, double* dx3, double* nrm3,  \
                                     double* om3, double* oml, double* eta3, double* dv3) {        \
    scene_##SUF* s = (scene_##SUF*)p;                                                              \
    fluid_##SUF* F = &s->fluid[f];                                                                 \
    for (size_t i = 0; i < s->n; ++i) {                                                            \
      if (lam) lam[i] = (double)F->lambda[i];                                                      \
      if (oml) oml[i] = (double)F->omlen[i];                                                       \
      for (int a = 0; a < 3; ++a) {                                                                \
        if (dx3) dx3[3 * i + a] = (double)F->dx[3 * i + a];                                        \
        if (nrm3) nrm3[3 * i + a] = (double)F->normal[3 * i + a];                                  \
        if (om3) om3[3 * i + a] = (double)F->omega[3 * i + a];                                     \
        if (eta3) eta3[3 * i + a] = (double)F->eta[3 * i + a];                                     \
        if (dv3) dv3[3 * i + a] = (double)F->dv[3 * i + a];                                        \
      }                                                                                            \
    }                                                                                              \
  }

ORC_SCENE_API(f32, float)
ORC_SCENE_API(f64, double)

/* SPH closed forms for the known-answer tests (kernels.ts:64-107). */
double orc_poly6_f32(double r2, double h) {
  return (double)poly6_f32((float)r2, (float)(h * h), (float)(315.0 / (64.0 * M_PI * pow(h, 9))));
}
void orc_spiky_f32(const double r[3], double h, double out[3]) {
  float rr[3] = {(float)r[0], (float)r[1], (float)r[2]}, g[3];
  spiky_f32(rr, (float)h, (float)(45.0 / (M_PI * pow(h, 6))), g);
  for (int a = 0; a < 3; ++a) out[a] = g[a];
}
