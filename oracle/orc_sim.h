/*
 * CPU restatement of one SimLoop substep schedule, included twice by
 * pf_oracle.c: with REAL=float (the FP32 mirror) and REAL=double (FP64).
 * TEST INFRASTRUCTURE ONLY: it checks the GPU path, it is never the product.
 *
 * Each function cites the reference pass it restates. Particles stay in
 * their original order; the neighbour order is the one a stable counting
 * sort gives (27 buckets dz->dy->dx, ascending particle index inside a
 * bucket), which is one of the orders the reference's atomic scatter may
 * produce (sort.ts:187-192).
 *
 * How each part is pinned (tests/test_oracle_*.py, tests/test_formats.py):
 *  - grid: the reference tests' collision-rate curve and Morton cells, the
 *    overflow rules, lattice counts;
 *  - neighbour list and fluid passes (tests/fluid_pins.py): the reference
 *    tests' brute-force neighbour sets, 64-entry cap and density sum; the
 *    walk order against an all-pairs search; lambda and position delta of a
 *    compressed pair in closed form and the point symmetry of a lattice's
 *    corrections; XSPH on a pair and a uniform drift; omega of a rigid
 *    rotation, eta and the confinement change of a linear |omega| field;
 *    the cohesion and adhesion splines' magnitudes on pairs; the viscosity
 *    solver's shear decay. Each plausible slip (a dropped factor, a sign, an
 *    operand order, a constant, <= for <) fails at least one of them
 *    (tests/test_oracle_planted_mistakes.py). What stays unpinned is the
 *    last-bit rounding of the reference's WGSL (its 2.5-ulp division and
 *    square root, its summation order), which cannot run here;
 *  - colliders: contact at exactly one radius per shape, moving shapes,
 *    dyadic fields sampled exactly, the half-float table;
 *  - boundaries, viscosity, contacts, shape matching: closed forms, a
 *    brute-force pair search, fixed points and the reference's scenarios;
 *  - surface field: a single particle's distance, the separable Gaussian,
 *    volume conservation, the motion-stretch centroid, plane proximity.
 *
 * Per-particle loops run in parallel (OpenMP) where each particle writes
 * only its own outputs; scatters into another particle's fixed-point
 * accumulator use atomic integer adds. Integer sums do not depend on order,
 * so every thread count gives the same bits.
 *
 * Expected macros: REAL, SFX(name) (symbol suffixing), R(x) (literal),
 * RSQRT, RFLOOR, RFMAX, RFMIN, RFABS, RSIN, RCOS, RRINT.
 */

typedef struct {
  uint32_t start, count, effects;
  REAL rho0, pv, mass, inv_mass, h, h2, poly6, spiky, compliance;
  REAL gamma, visc, vort, adhesion;
  REAL coh_coef, coh_h6_64, adh_coef, adh_m4_over_h, adh_2h;
  REAL imp_scale, imp_inv;
  REAL rea_scale, rea_inv;
  uint32_t nb;                                 /* boundary ranges (addBoundary) */
  uint32_t bstart[8], bcount[8], bdyn[8];
  uint32_t* nl;     /* [64 n] row-major: nl[i*64 + k] */
  uint32_t* nlc;    /* [n] */
  REAL* normal;     /* [3n] */
  REAL* lambda;     /* [n] */
  REAL* dx;         /* [3n] */
  REAL* omega;      /* [3n] */
  REAL* omlen;      /* [n] */
  REAL* eta;        /* [3n] */
  REAL* dv;         /* [3n] */
  double h_d;       /* smoothing radius in double (host-side constants) */
  uint32_t isv_iters;             /* ViscositySolver sweeps, 0: none */
  REAL isv_nu, isv_c10;           /* viscosity, 10 / h^2 */
  REAL *isv_rhs, *isv_a, *isv_b;  /* [3n] by particle id, zero outside the fluid */
} SFX(fluid);

typedef struct {
  uint32_t kind, flags;
  REAL d0[4], d1[4], vel[4], rot[4], spin[3];
} SFX(prim);

typedef struct {
  uint16_t* grid; /* half floats, x fastest */
  int nx, ny, nz;
  REAL res[3], inv_res[3], origin[3], voxel[3], inv_voxel[3], pos[3];
  REAL rot[9], inv_rot[9]; /* row-major */
  REAL scale, inv_scale, thickness, mus, muk;
} SFX(sdfc);

typedef struct {
  uint32_t n;
  REAL radius;
  REAL *x, *xs, *v, *w, *rho, *bvol; /* x, xs, v: [3n] */
  /* grid */
  REAL cell, origin[3];
  uint32_t H;
  uint32_t *key, *start, *perm; /* start: [H+1] */
  /* materials */
  uint32_t nfluids;
  SFX(fluid) fluid[4];
  uint32_t nprim;
  SFX(prim) prim[32];
  REAL* plam;      /* [2 * 32 * n] */
  uint32_t nsdf;
  SFX(sdfc) sdf[4];
  REAL* sdflam;    /* [2 * 4 * n] */
  int32_t *pacc, *vacc, *iacc, *racc; /* [3n] fixed-point accumulators */
  int initialized;                     /* first step ran (static boundary volumes) */
  REAL pos_scale, pos_inv, vel_scale, vel_inv;
  /* soft bodies (global) */
  uint32_t nbodies;
  uint32_t bstart[4096], bcount[4096];
  REAL bcomp[4096];
  REAL *rest, *weights, *slam; /* rest/slam: [3n] */
  REAL centers[4096][3];
  REAL rot[4096][9];
  /* local shape matching (shapeMatchImplicit.ts, rotation.ts), id order */
  int local;
  double radius_d;
  uint32_t nledges, capledges;
  uint32_t* ledges;            /* global-id pairs, bodies in order */
  uint32_t span_lo, span_hi;
  uint32_t *csr_off, *csr_idx; /* built on the first step */
  REAL *rcen, *gcen, *grot, *lcomp, *plam4; /* [3n], [3n], [9n], [n], [3 * entries] */
  int32_t* lacc;               /* [3n] */
  REAL *q, *qp, *om;           /* [4n] orientation, predicted orientation, angular velocity */
  /* particle contacts (src/core/contact), records in generation order */
  uint32_t* grp;                     /* [n] collision group */
  REAL* xs_grid;                     /* [3n] predictions at the last grid build (sortedPredictedPositions) */
  int contact;                       /* SimLoop contact option on */
  REAL ct_dist, ct_cr2, ct_mus, ct_muk; /* 2r, (2r * 1.1)^2, muS, muK */
  uint32_t ct_max, ct_count;         /* maxContacts; pairs emitted last substep (dropped included) */
  uint32_t nquiet, qstart[8], qcount[8]; /* extra noSelfContacts ranges (fluids are always quiet) */
  uint32_t *ct_i, *ct_j;             /* [ct_max] */
  int32_t *ct_ln, *ct_lt;            /* [ct_max] multipliers, 2^20 ticks per kg m */
  REAL* ct_n;                        /* [3 ct_max] last normal */
  /* loop */
  uint32_t substeps, iterations;
  REAL g[3], dt;
  REAL clock[64];
  /* flags */
  int flag_grid, flag_nl, flag_pos, flag_vel, flag_imp, flag_rea;
} SFX(scene);

/* --------------------------------------------------------------------- */
/* small vector helpers, operand order as the GPU kernels                */
/* --------------------------------------------------------------------- */
static inline REAL SFX(dot)(const REAL a[3], const REAL b[3]) {
  return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2];
}
static inline void SFX(cross)(const REAL a[3], const REAL b[3], REAL o[3]) {
  REAL t0 = a[1] * b[2] - a[2] * b[1];
  REAL t1 = a[2] * b[0] - a[0] * b[2];
  REAL t2 = a[0] * b[1] - a[1] * b[0];
  o[0] = t0; o[1] = t1; o[2] = t2;
}

/* kernels.ts:64-69 */
static inline REAL SFX(poly6)(REAL r2, REAL h2, REAL coef) {
  REAL d = RFMAX(h2 - r2, R(0.0));
  return ((d * d) * d) * coef;
}

/* kernels.ts:89-107: grad W = r * sp, with the scalar
 * sp = -(t*t*coef) / max(|r|, 1e-20), t = max(h - |r|, 0). The passes fold
 * their per-neighbour factor into sp before it meets r (one product per
 * component); the reference's WGSL divides per component with 2.5-ulp
 * division, and this stays within that. */
static inline REAL SFX(spiky_s)(const REAL r[3], REAL h, REAL coef) {
  REAL rsq = SFX(dot)(r, r);
  REAL rl = RSQRT(rsq);
  REAL t = RFMAX(h - rl, R(0.0));
  REAL inv = R(1.0) / RFMAX(rl, (REAL)1e-20);
  REAL mag = -((t * t) * coef);
  return inv * mag;
}
static inline void SFX(spiky)(const REAL r[3], REAL h, REAL coef, REAL g[3]) {
  REAL sp = SFX(spiky_s)(r, h, coef);
  g[0] = r[0] * sp;
  g[1] = r[1] * sp;
  g[2] = r[2] * sp;
}

/* accumulator.ts:62-68: i32(round(v * scale)), half-to-even, saturating */
static inline int32_t SFX(ticks)(REAL v, REAL scale) {
  REAL t = RRINT(v * scale);
  if (!(t == t)) return 0;
  if (t >= (REAL)2147483648.0) return 2147483647;
  if (t < (REAL)-2147483648.0) return (int32_t)(-2147483647 - 1);
  return (int32_t)t;
}
static inline int32_t SFX(wadd)(int32_t a, int32_t b) { return (int32_t)((uint32_t)a + (uint32_t)b); }

/* accumulator.ts:101-131: returns 1 when saturated */
static inline int SFX(acc_change)(const int32_t* t, REAL inv, REAL c[3]) {
  REAL tx = (REAL)t[0], ty = (REAL)t[1], tz = (REAL)t[2];
  REAL m = RFMAX(RFMAX(RFABS(tx), RFABS(ty)), RFABS(tz));
  c[0] = tx * inv; c[1] = ty * inv; c[2] = tz * inv;
  return m >= (REAL)1073741824.0;
}

/* --------------------------------------------------------------------- */
/* grid (cellIndex.ts:113-170, sort.ts, stable)                           */
/* --------------------------------------------------------------------- */
static int32_t SFX(bin_coord)(REAL p, REAL o, REAL cell, int* over) {
  REAL rel = (p - o) / cell;
  int32_t raw = orc_f2i_sat((double)RFLOOR(rel));
  int32_t c = raw < -(1 << 30) ? -(1 << 30) : (raw > (1 << 30) - 1 ? (1 << 30) - 1 : raw);
  if (raw != c || c < -512 || c >= 512) *over = 1;
  return c;
}

static void SFX(grid_build)(SFX(scene)* s) {
  uint32_t n = s->n, H = s->H;
  uint32_t* counts = (uint32_t*)calloc(H, sizeof(uint32_t));
  int over = 0;
  for (uint32_t i = 0; i < n; ++i) {
    const REAL* p = s->xs + 3 * (size_t)i;
    int32_t cx = SFX(bin_coord)(p[0], s->origin[0], s->cell, &over);
    int32_t cy = SFX(bin_coord)(p[1], s->origin[1], s->cell, &over);
    int32_t cz = SFX(bin_coord)(p[2], s->origin[2], s->cell, &over);
    uint32_t b = orc_morton(cx, cy, cz) & (H - 1);
    s->key[i] = b;
    counts[b]++;
  }
  uint32_t run = 0;
  for (uint32_t b = 0; b < H; ++b) {
    s->start[b] = run;
    run += counts[b];
  }
  s->start[H] = run;
  for (uint32_t b = 0; b < H; ++b) counts[b] = s->start[b];
  for (uint32_t i = 0; i < n; ++i) s->perm[counts[s->key[i]]++] = i;
  free(counts);
  s->flag_grid = over;
  /* the grid's sorted copy of the predictions (sortedPositions.ts:31-39) */
  memcpy(s->xs_grid, s->xs, sizeof(REAL) * 3 * (size_t)n);
}

/* neighborList.ts:45-72 over query.ts:24-70: the query point is the current
 * prediction, the candidates' positions come from the grid's sorted copy
 * (neighborList.ts:59); the two differ only where a contact stabilisation
 * moved particles after the grid build. */
static void SFX(neighbors)(SFX(scene)* s, SFX(fluid)* F) {
  uint32_t H = s->H;
  int over = 0;
#pragma omp parallel for schedule(dynamic, 128) reduction(|| : over)
  for (uint32_t i = F->start; i < F->start + F->count; ++i) {
    const REAL* xi = s->xs + 3 * (size_t)i;
    int32_t q[3];
    for (int a = 0; a < 3; ++a) q[a] = orc_f2i_sat((double)RFLOOR((xi[a] - s->origin[a]) / s->cell));
    uint32_t visited[27], nv = 0, cnt = 0;
    for (int dz = -1; dz <= 1; ++dz)
      for (int dy = -1; dy <= 1; ++dy)
        for (int dx = -1; dx <= 1; ++dx) {
          uint32_t b = orc_morton((int32_t)((uint32_t)q[0] + (uint32_t)dx), (int32_t)((uint32_t)q[1] + (uint32_t)dy),
                                  (int32_t)((uint32_t)q[2] + (uint32_t)dz)) & (H - 1);
          int seen = 0;
          for (uint32_t k = 0; k < nv; ++k) seen |= visited[k] == b;
          visited[nv++] = b;
          if (seen) continue;
          for (uint32_t t = s->start[b]; t < s->start[b + 1]; ++t) {
            uint32_t j = s->perm[t];
            const REAL* xj = s->xs_grid + 3 * (size_t)j;
            REAL d[3] = {xi[0] - xj[0], xi[1] - xj[1], xi[2] - xj[2]};
            if (SFX(dot)(d, d) < F->h2) {
              if (cnt < 64) F->nl[(size_t)i * 64 + cnt++] = j;
              else over = 1;
            }
          }
        }
    F->nlc[i] = cnt;
  }
  s->flag_nl |= over;
}

static inline REAL SFX(nvol)(const SFX(scene)* s, const SFX(fluid)* F, uint32_t j) {
  return s->bvol[j] > R(0.0) ? s->bvol[j] : F->pv;
}
static inline REAL SFX(nmass)(const SFX(scene)* s, const SFX(fluid)* F, uint32_t j) {
  if (s->bvol[j] > R(0.0)) return F->rho0 * s->bvol[j];
  return s->w[j] > R(0.0) ? R(1.0) / s->w[j] : R(0.0);
}

/* boundaryVolume.ts:26-52: V_i = 1 / max(sum_k W(x_i - x_k), 1e-12) over the
 * boundary's own range, walking the grid (query.ts:24-70, bucket dedup).
 * pos: the positions the walk reads (committed for a static boundary's
 * one-time pass, predicted for a dynamic one). */
static void SFX(boundary_volume)(SFX(scene)* s, const SFX(fluid)* F, uint32_t b, const REAL* pos) {
  uint32_t H = s->H, lo = F->bstart[b], hi = F->bstart[b] + F->bcount[b];
#pragma omp parallel for schedule(dynamic, 128)
  for (uint32_t i = lo; i < hi; ++i) {
    const REAL* xi = pos + 3 * (size_t)i;
    int32_t q[3];
    for (int a = 0; a < 3; ++a) q[a] = orc_f2i_sat((double)RFLOOR((xi[a] - s->origin[a]) / s->cell));
    uint32_t visited[27], nv = 0;
    REAL sum = 0;
    for (int dz = -1; dz <= 1; ++dz)
      for (int dy = -1; dy <= 1; ++dy)
        for (int dx = -1; dx <= 1; ++dx) {
          uint32_t bk = orc_morton((int32_t)((uint32_t)q[0] + (uint32_t)dx), (int32_t)((uint32_t)q[1] + (uint32_t)dy),
                                   (int32_t)((uint32_t)q[2] + (uint32_t)dz)) & (H - 1);
          int seen = 0;
          for (uint32_t k = 0; k < nv; ++k) seen |= visited[k] == bk;
          visited[nv++] = bk;
          if (seen) continue;
          for (uint32_t t = s->start[bk]; t < s->start[bk + 1]; ++t) {
            uint32_t j = s->perm[t];
            if (j < lo || j >= hi) continue;
            const REAL* xj = pos + 3 * (size_t)j;
            REAL d[3] = {xi[0] - xj[0], xi[1] - xj[1], xi[2] - xj[2]};
            REAL r2 = SFX(dot)(d, d);
            if (r2 >= F->h2) continue;
            sum = sum + SFX(poly6)(r2, F->h2, F->poly6);
          }
        }
    s->bvol[i] = R(1.0) / RFMAX(sum, (REAL)1e-12);
  }
}

static inline int SFX(in_boundary)(const SFX(fluid)* F, uint32_t j) {
  for (uint32_t b = 0; b < F->nb; ++b)
    if (j >= F->bstart[b] && j < F->bstart[b] + F->bcount[b]) return 1;
  return 0;
}

/* cohesion.ts:25-39 */
static void SFX(color_normal)(SFX(scene)* s, SFX(fluid)* F) {
#pragma omp parallel for schedule(dynamic, 128)
  for (uint32_t i = F->start; i < F->start + F->count; ++i) {
    const REAL* xi = s->xs + 3 * (size_t)i;
    REAL sum[3] = {0, 0, 0};
    for (uint32_t k = 0; k < F->nlc[i]; ++k) {
      uint32_t j = F->nl[(size_t)i * 64 + k];
      const REAL* xj = s->xs + 3 * (size_t)j;
      REAL d[3] = {xi[0] - xj[0], xi[1] - xj[1], xi[2] - xj[2]};
      REAL V = SFX(nvol)(s, F, j);
      REAL sc = SFX(spiky_s)(d, F->h, F->spiky) * V;
      sum[0] = sum[0] + d[0] * sc; sum[1] = sum[1] + d[1] * sc; sum[2] = sum[2] + d[2] * sc;
    }
    for (int a = 0; a < 3; ++a) F->normal[3 * (size_t)i + a] = sum[a] * F->h;
  }
}

/* cohesion.ts:46-55 */
static inline REAL SFX(coh_spline)(const SFX(fluid)* F, REAL r) {
  REAL gap = RFMAX(F->h - r, R(0.0));
  REAL shared = ((gap * gap) * gap) * ((r * r) * r);
  REAL value = (r * R(2.0) > F->h) ? shared : (shared * R(2.0) - F->coh_h6_64);
  return (r > R(0.0) && r <= F->h) ? value * F->coh_coef : R(0.0);
}

/* adhesion.ts:16-27 */
static inline REAL SFX(adh_spline)(const SFX(fluid)* F, REAL r) {
  REAL root = RSQRT(RSQRT(RFMAX(((r * r) * F->adh_m4_over_h + r * R(6.0)) - F->adh_2h, R(0.0))));
  return (r * R(2.0) > F->h && r <= F->h) ? root * F->adh_coef : R(0.0);
}

/* cohesion.ts:63-103 (pairs j > i inside the range, scattered both ways),
 * adhesion.ts:34-60, then the impulse apply (FluidSystem.ts:253-258). */
static void SFX(impulses)(SFX(scene)* s, SFX(fluid)* F) {
  int st = (F->effects & 4u) != 0, adh = (F->effects & 8u) != 0;
  uint32_t end = F->start + F->count;
#pragma omp parallel for schedule(dynamic, 128)
  for (uint32_t i = F->start; i < end; ++i) {
    const REAL* xi = s->xs + 3 * (size_t)i;
    for (uint32_t k = 0; k < F->nlc[i]; ++k) {
      uint32_t j = F->nl[(size_t)i * 64 + k];
      const REAL* xj = s->xs + 3 * (size_t)j;
      if (st && j > i && j < end) {
        REAL off[3] = {xi[0] - xj[0], xi[1] - xj[1], xi[2] - xj[2]};
        REAL r = RSQRT(SFX(dot)(off, off));
        if (!(r < (REAL)1e-20)) {
          REAL corr = (F->rho0 * R(2.0)) / RFMAX(s->rho[i] + s->rho[j], R(1.0));
          REAL cs = -(((F->gamma * F->mass) * F->mass) * SFX(coh_spline)(F, r));
          REAL gm = F->gamma * F->mass;
          REAL ir = R(1.0) / r;
          for (int a = 0; a < 3; ++a) {
            REAL coh = (off[a] * ir) * cs;
            REAL curv = -((F->normal[3 * (size_t)i + a] - F->normal[3 * (size_t)j + a]) * gm);
            REAL dv = (((coh + curv) * corr) * F->inv_mass) * s->dt;
            orc_acc_add(&s->iacc[3 * (size_t)i + a], SFX(ticks)(dv, F->imp_scale));
            orc_acc_add(&s->iacc[3 * (size_t)j + a], SFX(ticks)(-dv, F->imp_scale));
          }
        }
      }
      if (adh && s->bvol[j] > R(0.0)) {
        REAL off[3] = {xi[0] - xj[0], xi[1] - xj[1], xi[2] - xj[2]};
        REAL r = RSQRT(SFX(dot)(off, off));
        if (!(r < (REAL)1e-20)) {
          REAL fs = -(((F->adhesion * F->mass) * (F->rho0 * s->bvol[j])) * SFX(adh_spline)(F, r));
          REAL ir = R(1.0) / r;
          for (int a = 0; a < 3; ++a) {
            REAL f = (off[a] * ir) * fs;
            orc_acc_add(&s->iacc[3 * (size_t)i + a], SFX(ticks)((f * F->inv_mass) * s->dt, F->imp_scale));
          }
        }
      }
    }
  }
  int sat = 0;
#pragma omp parallel for schedule(static) reduction(|| : sat)
  for (uint32_t i = F->start; i < end; ++i) {
    REAL c[3];
    if (SFX(acc_change)(s->iacc + 3 * (size_t)i, F->imp_inv, c)) sat = 1;
    for (int a = 0; a < 3; ++a) {
      s->v[3 * (size_t)i + a] = s->v[3 * (size_t)i + a] + c[a];
      s->xs[3 * (size_t)i + a] = s->xs[3 * (size_t)i + a] + c[a] * s->dt;
      s->iacc[3 * (size_t)i + a] = 0;
    }
  }
  if (sat) s->flag_imp = 1;
}

/* lambda.ts:30-71 */
static void SFX(lambda_pass)(SFX(scene)* s, SFX(fluid)* F) {
#pragma omp parallel for schedule(dynamic, 128)
  for (uint32_t i = F->start; i < F->start + F->count; ++i) {
    const REAL* xi = s->xs + 3 * (size_t)i;
    REAL rho = 0, gss = 0, gs[3] = {0, 0, 0};
    for (uint32_t k = 0; k < F->nlc[i]; ++k) {
      uint32_t j = F->nl[(size_t)i * 64 + k];
      const REAL* xj = s->xs + 3 * (size_t)j;
      REAL d[3] = {xi[0] - xj[0], xi[1] - xj[1], xi[2] - xj[2]}, g[3];
      REAL mj = SFX(nmass)(s, F, j);
      rho = rho + mj * SFX(poly6)(SFX(dot)(d, d), F->h2, F->poly6);
      REAL sc = SFX(spiky_s)(d, F->h, F->spiky) * mj;
      g[0] = d[0] * sc; g[1] = d[1] * sc; g[2] = d[2] * sc;
      gs[0] = gs[0] + g[0]; gs[1] = gs[1] + g[1]; gs[2] = gs[2] + g[2];
      /* |g|^2 = |d|^2 sc^2 */
      gss = gss + ((SFX(dot)(d, d) * sc) * sc) * s->w[j];
    }
    s->rho[i] = rho;
    REAL inv = R(1.0) / F->rho0;
    REAL C = rho * inv - R(1.0);
    REAL lam = 0;
    if (!(C <= R(0.0))) {
      REAL gn = (SFX(dot)(gs, gs) * s->w[i] + gss) * (inv * inv);
      REAL at = F->compliance / (s->dt * s->dt);
      lam = -C / RFMAX(gn + at, (REAL)1e-20);
    }
    F->lambda[i] = lam;
  }
}

/* positionDelta.ts:26-85 (deltaX pass, then the separate apply) */
static void SFX(delta_pass)(SFX(scene)* s, SFX(fluid)* F) {
  uint32_t end = F->start + F->count;
#pragma omp parallel for schedule(dynamic, 128)
  for (uint32_t i = F->start; i < end; ++i) {
    const REAL* xi = s->xs + 3 * (size_t)i;
    REAL li = F->lambda[i], sum[3] = {0, 0, 0};
    for (uint32_t k = 0; k < F->nlc[i]; ++k) {
      uint32_t j = F->nl[(size_t)i * 64 + k];
      const REAL* xj = s->xs + 3 * (size_t)j;
      REAL d[3] = {xi[0] - xj[0], xi[1] - xj[1], xi[2] - xj[2]}, g[3];
      REAL sp = SFX(spiky_s)(d, F->h, F->spiky);
      REAL lj = (j >= F->start && j < end) ? F->lambda[j] : R(0.0);
      REAL sc = (li + lj) * SFX(nmass)(s, F, j);
      REAL gsc = sp * sc;
      sum[0] = sum[0] + d[0] * gsc; sum[1] = sum[1] + d[1] * gsc; sum[2] = sum[2] + d[2] * gsc;
      g[0] = d[0] * sp; g[1] = d[1] * sp; g[2] = d[2] * sp;
      /* reaction on dynamic boundary neighbours (positionDelta.ts:54-69) */
      if (F->nb && s->bvol[j] > R(0.0) && s->w[j] > R(0.0) && li != R(0.0) && SFX(in_boundary)(F, j)) {
        REAL rs = ((-(li / F->rho0)) * (F->rho0 * s->bvol[j])) * s->w[j];
        for (int a = 0; a < 3; ++a)
          orc_acc_add(&s->racc[3 * (size_t)j + a], SFX(ticks)(g[a] * rs, F->rea_scale));
      }
    }
    for (int a = 0; a < 3; ++a) F->dx[3 * (size_t)i + a] = (sum[a] / F->rho0) * s->w[i];
  }
#pragma omp parallel for schedule(static)
  for (uint32_t i = F->start; i < end; ++i)
    for (int a = 0; a < 3; ++a) s->xs[3 * (size_t)i + a] = s->xs[3 * (size_t)i + a] + F->dx[3 * (size_t)i + a];
  if (F->nb) {
    /* reaction apply over the covering range, then clear (accumulator.ts:101-131) */
    uint32_t lo = 0xffffffffu, hi = 0;
    for (uint32_t b = 0; b < F->nb; ++b) {
      if (F->bstart[b] < lo) lo = F->bstart[b];
      if (F->bstart[b] + F->bcount[b] > hi) hi = F->bstart[b] + F->bcount[b];
    }
    for (uint32_t j = lo; j < hi; ++j) {
      REAL c[3];
      s->flag_rea |= SFX(acc_change)(s->racc + 3 * (size_t)j, F->rea_inv, c);
      for (int a = 0; a < 3; ++a) {
        s->xs[3 * (size_t)j + a] = s->xs[3 * (size_t)j + a] + c[a];
        s->racc[3 * (size_t)j + a] = 0;
      }
    }
  }
}

/* velocity.ts:36-119 */
static void SFX(velocity_passes)(SFX(scene)* s, SFX(fluid)* F) {
  int vort = (F->effects & 2u) != 0, visc = (F->effects & 1u) != 0;
  if (!vort && !visc) return;
  uint32_t end = F->start + F->count;
#pragma omp parallel for schedule(dynamic, 128)
  for (uint32_t i = F->start; i < end; ++i) {
    const REAL* xi = s->x + 3 * (size_t)i;
    const REAL* vi = s->v + 3 * (size_t)i;
    REAL curl[3] = {0, 0, 0}, sm[3] = {0, 0, 0};
    for (uint32_t k = 0; k < F->nlc[i]; ++k) {
      uint32_t j = F->nl[(size_t)i * 64 + k];
      const REAL* xj = s->x + 3 * (size_t)j;
      const REAL* vj = s->v + 3 * (size_t)j;
      REAL off[3] = {xi[0] - xj[0], xi[1] - xj[1], xi[2] - xj[2]};
      REAL V = SFX(nvol)(s, F, j);
      if (vort) {
        REAL g[3], dv[3] = {vi[0] - vj[0], vi[1] - vj[1], vi[2] - vj[2]}, c[3];
        REAL sc = SFX(spiky_s)(off, F->h, F->spiky) * V;
        g[0] = off[0] * sc; g[1] = off[1] * sc; g[2] = off[2] * sc;
        SFX(cross)(dv, g, c);
        curl[0] = curl[0] + c[0]; curl[1] = curl[1] + c[1]; curl[2] = curl[2] + c[2];
      }
      if (visc) {
        REAL W = SFX(poly6)(SFX(dot)(off, off), F->h2, F->poly6);
        for (int a = 0; a < 3; ++a) sm[a] = sm[a] + ((vj[a] - vi[a]) * W) * V;
      }
    }
    if (vort) {
      for (int a = 0; a < 3; ++a) F->omega[3 * (size_t)i + a] = curl[a];
      F->omlen[i] = RSQRT(SFX(dot)(curl, curl));
    }
    if (visc) for (int a = 0; a < 3; ++a) F->dv[3 * (size_t)i + a] = sm[a] * F->visc;
  }
  if (vort) {
#pragma omp parallel for schedule(dynamic, 128)
    for (uint32_t i = F->start; i < end; ++i) {
      const REAL* xi = s->x + 3 * (size_t)i;
      REAL oi = F->omlen[i], eta[3] = {0, 0, 0};
      for (uint32_t k = 0; k < F->nlc[i]; ++k) {
        uint32_t j = F->nl[(size_t)i * 64 + k];
        if (s->bvol[j] > R(0.0)) continue;
        const REAL* xj = s->x + 3 * (size_t)j;
        REAL off[3] = {xi[0] - xj[0], xi[1] - xj[1], xi[2] - xj[2]};
        REAL oj = (j >= F->start && j < end) ? F->omlen[j] : R(0.0);
        REAL sc = (oj - oi) * F->pv;
        REAL gsc = SFX(spiky_s)(off, F->h, F->spiky) * sc;
        eta[0] = eta[0] + off[0] * gsc; eta[1] = eta[1] + off[1] * gsc; eta[2] = eta[2] + off[2] * gsc;
      }
      for (int a = 0; a < 3; ++a) F->eta[3 * (size_t)i + a] = eta[a];
    }
  }
#pragma omp parallel for schedule(static)
  for (uint32_t i = F->start; i < end; ++i) {
    REAL ch[3] = {0, 0, 0};
    if (vort) {
      REAL* eta = F->eta + 3 * (size_t)i;
      REAL el = RSQRT(SFX(dot)(eta, eta));
      if (el >= (REAL)1e-12) {
        REAL nrm[3] = {eta[0] / el, eta[1] / el, eta[2] / el}, f[3];
        SFX(cross)(nrm, F->omega + 3 * (size_t)i, f);
        for (int a = 0; a < 3; ++a) ch[a] = ch[a] + (f[a] * F->vort) * s->dt;
      }
    }
    if (visc) for (int a = 0; a < 3; ++a) ch[a] = ch[a] + F->dv[3 * (size_t)i + a];
    for (int a = 0; a < 3; ++a) s->v[3 * (size_t)i + a] = s->v[3 * (size_t)i + a] + ch[a];
  }
}

/* ViscositySolver (ViscositySolver.ts:78-123): seed rhs and the first
 * iterate with the fluid's velocities, Jacobi sweeps of
 * (I - dt nu L) v = v0 over the neighbour list, then copy the last iterate
 * into the velocities. Both iterates are zero outside the fluid, so other
 * neighbours add weight but no velocity. */
static void SFX(viscosity_solve)(SFX(scene)* s, SFX(fluid)* F) {
  uint32_t end = F->start + F->count;
  REAL* rhs = F->isv_rhs;
  for (uint32_t i = F->start; i < end; ++i)
    for (int a = 0; a < 3; ++a) {
      rhs[3 * (size_t)i + a] = s->v[3 * (size_t)i + a];
      F->isv_a[3 * (size_t)i + a] = s->v[3 * (size_t)i + a];
    }
  REAL rate = (F->isv_nu * s->dt) * F->isv_c10;
  for (uint32_t k = 0; k < F->isv_iters; ++k) {
    const REAL* src = (k & 1u) ? F->isv_b : F->isv_a;
    REAL* dst = (k & 1u) ? F->isv_a : F->isv_b;
#pragma omp parallel for schedule(dynamic, 128)
    for (uint32_t i = F->start; i < end; ++i) {
      const REAL* xi = s->x + 3 * (size_t)i;
      REAL sum[3] = {0, 0, 0}, weight = 0;
      for (uint32_t k2 = 0; k2 < F->nlc[i]; ++k2) {
        uint32_t j = F->nl[(size_t)i * 64 + k2];
        const REAL* xj = s->x + 3 * (size_t)j;
        REAL off[3] = {xi[0] - xj[0], xi[1] - xj[1], xi[2] - xj[2]};
        REAL w = SFX(poly6)(SFX(dot)(off, off), F->h2, F->poly6) * F->pv;
        for (int a = 0; a < 3; ++a) sum[a] = sum[a] + src[3 * (size_t)j + a] * w;
        weight = weight + w;
      }
      for (int a = 0; a < 3; ++a)
        dst[3 * (size_t)i + a] = (rhs[3 * (size_t)i + a] + sum[a] * rate) / (weight * rate + R(1.0));
    }
  }
  const REAL* res = (F->isv_iters & 1u) ? F->isv_b : F->isv_a;
  for (uint32_t i = F->start; i < end; ++i)
    for (int a = 0; a < 3; ++a) s->v[3 * (size_t)i + a] = res[3 * (size_t)i + a];
}

/* --------------------------------------------------------------------- */
/* colliders (primitives.ts:85-242, collider.ts:53-118)                  */
/* --------------------------------------------------------------------- */
static void SFX(qrot)(const REAL q[4], const REAL v[3], REAL o[3]) {
  REAL c[3], t[3], c2[3];
  SFX(cross)(q, v, c);
  t[0] = R(2.0) * c[0]; t[1] = R(2.0) * c[1]; t[2] = R(2.0) * c[2];
  SFX(cross)(q, t, c2);
  for (int a = 0; a < 3; ++a) o[a] = (v[a] + t[a] * q[3]) + c2[a];
}

static void SFX(pivot)(const SFX(prim)* P, REAL o[3]) {
  for (int a = 0; a < 3; ++a) {
    if (P->kind == 0) o[a] = P->d1[a];
    else if (P->kind == 3) o[a] = (P->d0[a] + P->d1[a]) * R(0.5);
    else o[a] = P->d0[a];
  }
}

/* The primitive's signed distance and gradient at x as it stands now
 * (primitives.ts:119-217), inversion included. */
static void SFX(prim_shape)(const SFX(prim)* P, const REAL x[3], REAL* phi_out, REAL g_out[3]);

static void SFX(sdf)(const SFX(prim)* P, const REAL xw[3], REAL rewind, REAL* phi_out, REAL grad[3]) {
  REAL t = (P->flags & 2u) ? rewind : R(0.0);
  REAL shifted[3], pv[3], x[3], turn[4], rel[3], rr[3];
  for (int a = 0; a < 3; ++a) shifted[a] = xw[a] + P->vel[a] * t;
  SFX(pivot)(P, pv);
  {
    REAL rate = RSQRT(SFX(dot)(P->spin, P->spin));
    REAL half = (rate * t) * R(0.5);
    REAL den = RFMAX(rate, (REAL)1e-8);
    REAL sn = RSIN(half);
    for (int a = 0; a < 3; ++a) turn[a] = (P->spin[a] / den) * sn;
    turn[3] = RCOS(half);
  }
  for (int a = 0; a < 3; ++a) rel[a] = shifted[a] - pv[a];
  SFX(qrot)(turn, rel, rr);
  for (int a = 0; a < 3; ++a) x[a] = pv[a] + rr[a];
  REAL phi, g[3];
  SFX(prim_shape)(P, x, &phi, g);
  *phi_out = phi;
  REAL un[4] = {-turn[0], -turn[1], -turn[2], turn[3]};
  SFX(qrot)(un, g, grad);
}

static void SFX(prim_shape)(const SFX(prim)* P, const REAL x[3], REAL* phi_out, REAL g_out[3]) {
  REAL phi = 0, g[3] = {0, 0, 0};
  if (P->kind == 0) {
    REAL d[3] = {x[0] - P->d1[0], x[1] - P->d1[1], x[2] - P->d1[2]};
    phi = SFX(dot)(P->d0, d);
    g[0] = P->d0[0]; g[1] = P->d0[1]; g[2] = P->d0[2];
  } else if (P->kind == 1) {
    REAL d[3] = {x[0] - P->d0[0], x[1] - P->d0[1], x[2] - P->d0[2]};
    REAL len = RSQRT(SFX(dot)(d, d));
    phi = len - P->d0[3];
    REAL sl = RFMAX(len, (REAL)1e-8);
    for (int a = 0; a < 3; ++a) g[a] = d[a] / sl;
  } else if (P->kind == 2) {
    REAL conj[4] = {-P->rot[0], -P->rot[1], -P->rot[2], P->rot[3]};
    REAL off[3] = {x[0] - P->d0[0], x[1] - P->d0[1], x[2] - P->d0[2]}, xl[3], q[3], qp[3], sg[3], lg[3];
    SFX(qrot)(conj, off, xl);
    for (int a = 0; a < 3; ++a) {
      q[a] = RFABS(xl[a]) - P->d1[a];
      qp[a] = RFMAX(q[a], R(0.0));
      sg[a] = xl[a] >= R(0.0) ? R(1.0) : R(-1.0);
    }
    REAL outside = RSQRT(SFX(dot)(qp, qp));
    REAL inside = RFMIN(RFMAX(RFMAX(q[0], q[1]), q[2]), R(0.0));
    phi = outside + inside;
    int domx = q[0] >= q[1] && q[0] >= q[2];
    int domy = q[1] >= q[0] && q[1] >= q[2];
    if (outside > (REAL)1e-8) {
      REAL ol = RFMAX(outside, (REAL)1e-8);
      for (int a = 0; a < 3; ++a) lg[a] = (qp[a] * sg[a]) / ol;
    } else {
      lg[0] = domx ? sg[0] : R(0.0);
      lg[1] = (!domx && domy) ? sg[1] : R(0.0);
      lg[2] = (!domx && !domy) ? sg[2] : R(0.0);
    }
    SFX(qrot)(P->rot, lg, g);
  } else {
    REAL ab[3], ax[3], cl[3], d[3];
    for (int a = 0; a < 3; ++a) { ab[a] = P->d1[a] - P->d0[a]; ax[a] = x[a] - P->d0[a]; }
    REAL ll = RFMAX(SFX(dot)(ab, ab), (REAL)1e-8);
    REAL tt = RFMIN(RFMAX(SFX(dot)(ax, ab) / ll, R(0.0)), R(1.0));
    for (int a = 0; a < 3; ++a) { cl[a] = P->d0[a] + ab[a] * tt; d[a] = x[a] - cl[a]; }
    REAL len = RSQRT(SFX(dot)(d, d));
    phi = len - P->d0[3];
    REAL sl = RFMAX(len, (REAL)1e-8);
    for (int a = 0; a < 3; ++a) g[a] = d[a] / sl;
  }
  if (P->flags & 1u) {
    phi = -phi;
    g[0] = -g[0]; g[1] = -g[1]; g[2] = -g[2];
  }
  *phi_out = phi;
  g_out[0] = g[0]; g_out[1] = g[1]; g_out[2] = g[2];
}

static void SFX(surface_velocity)(const SFX(prim)* P, const REAL xw[3], REAL rewind, REAL o[3]) {
  REAL pv[3], arm[3], c[3];
  SFX(pivot)(P, pv);
  REAL t = (P->flags & 2u) ? rewind : R(0.0);
  for (int a = 0; a < 3; ++a) arm[a] = (xw[a] - pv[a]) + P->vel[a] * t;
  SFX(cross)(P->spin, arm, c);
  for (int a = 0; a < 3; ++a) o[a] = P->vel[a] + c[a];
}

/* PrimitiveSet.ts:336-364 then the positions accumulator apply (loop.ts:222) */
/* half-float voxel (i, j, k) */
static inline REAL SFX(texel)(const SFX(sdfc)* D, int i, int j, int k) {
  return (REAL)orc_half_to_double(D->grid[((size_t)k * D->ny + j) * D->nx + i]);
}
static inline REAL SFX(lerp)(REAL a, REAL b, REAL f) { return a + (b - a) * f; }
static inline int SFX(clampi)(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

/* Linear filtering at normalised uv with clamp-to-edge addressing: texel
 * coordinate t = uv * n - 0.5, weights from its fraction, lerp x, y, z. */
static REAL SFX(tex)(const SFX(sdfc)* D, REAL u, REAL v, REAL w) {
  REAL t[3] = {u * D->res[0] - R(0.5), v * D->res[1] - R(0.5), w * D->res[2] - R(0.5)};
  int n[3] = {D->nx, D->ny, D->nz}, i0[3], i1[3];
  REAL f[3];
  for (int a = 0; a < 3; ++a) {
    REAL fl = RFLOOR(t[a]);
    f[a] = t[a] - fl;
    int ii = SFX(clampi)(orc_f2i_sat((double)fl), -1, n[a] - 1);
    i0[a] = ii < 0 ? 0 : ii;
    i1[a] = ii + 1 > n[a] - 1 ? n[a] - 1 : ii + 1;
  }
  REAL c00 = SFX(lerp)(SFX(texel)(D, i0[0], i0[1], i0[2]), SFX(texel)(D, i1[0], i0[1], i0[2]), f[0]);
  REAL c10 = SFX(lerp)(SFX(texel)(D, i0[0], i1[1], i0[2]), SFX(texel)(D, i1[0], i1[1], i0[2]), f[0]);
  REAL c01 = SFX(lerp)(SFX(texel)(D, i0[0], i0[1], i1[2]), SFX(texel)(D, i1[0], i0[1], i1[2]), f[0]);
  REAL c11 = SFX(lerp)(SFX(texel)(D, i0[0], i1[1], i1[2]), SFX(texel)(D, i1[0], i1[1], i1[2]), f[0]);
  return SFX(lerp)(SFX(lerp)(c00, c10, f[1]), SFX(lerp)(c01, c11, f[1]), f[2]);
}

/* emitSampleSdf (sdf.ts:107-146): world -> local -> uv, trilinear phi,
 * 6-tap central-difference gradient, back to world. */
static void SFX(sdf_sample)(const SFX(sdfc)* D, const REAL x[3], REAL* phi, REAL g[3]) {
  REAL wo[3] = {x[0] - D->pos[0], x[1] - D->pos[1], x[2] - D->pos[2]}, uv[3];
  for (int r = 0; r < 3; ++r) {
    REAL xl = ((D->inv_rot[3 * r] * wo[0] + D->inv_rot[3 * r + 1] * wo[1]) + D->inv_rot[3 * r + 2] * wo[2]) * D->inv_scale;
    uv[r] = ((xl - D->origin[r]) * D->inv_voxel[r]) / D->res[r];
  }
  REAL pl = SFX(tex)(D, uv[0], uv[1], uv[2]);
  REAL gl[3];
  gl[0] = (SFX(tex)(D, uv[0] + D->inv_res[0], uv[1], uv[2]) - SFX(tex)(D, uv[0] - D->inv_res[0], uv[1], uv[2])) / (R(2.0) * D->voxel[0]);
  gl[1] = (SFX(tex)(D, uv[0], uv[1] + D->inv_res[1], uv[2]) - SFX(tex)(D, uv[0], uv[1] - D->inv_res[1], uv[2])) / (R(2.0) * D->voxel[1]);
  gl[2] = (SFX(tex)(D, uv[0], uv[1], uv[2] + D->inv_res[2]) - SFX(tex)(D, uv[0], uv[1], uv[2] - D->inv_res[2])) / (R(2.0) * D->voxel[2]);
  *phi = pl * D->scale;
  for (int r = 0; r < 3; ++r) g[r] = (D->rot[3 * r] * gl[0] + D->rot[3 * r + 1] * gl[1]) + D->rot[3 * r + 2] * gl[2];
}

/* SDFCollider solve (SDFCollider.ts:237-269 + collider.ts:53-89), static
 * collider: ticks into the positions accumulator of particle p. */
static void SFX(sdf_solve)(SFX(scene)* s, uint32_t p, REAL w, const REAL* xs, const REAL disp[3]) {
  for (uint32_t k = 0; k < s->nsdf; ++k) {
    const SFX(sdfc)* D = &s->sdf[k];
    REAL phi, g[3];
    SFX(sdf_sample)(D, xs, &phi, g);
    REAL gl = RSQRT(SFX(dot)(g, g));
    if (!(gl > (REAL)1e-6)) continue;
    REAL nrm[3] = {g[0] / gl, g[1] / gl, g[2] / gl};
    REAL reach = s->radius + D->thickness;
    if (!(phi < reach)) continue;
    REAL depth = reach - phi;
    REAL* lam = s->sdflam + ((size_t)p * 4 + k) * 2;
    REAL ln = lam[0] + depth / w;
    lam[0] = ln;
    REAL sd = SFX(dot)(disp, nrm), tang[3];
    for (int a = 0; a < 3; ++a) tang[a] = disp[a] - nrm[a] * sd;
    REAL tl = RSQRT(SFX(dot)(tang, tang));
    REAL lt = lam[1];
    int sticks = tl > (REAL)1e-10 && (tl / w) <= (D->mus * ln - lt);
    for (int a = 0; a < 3; ++a) {
      REAL push = nrm[a] * depth;
      REAL corr = sticks ? push - tang[a] : push;
      s->pacc[3 * (size_t)p + a] = SFX(wadd)(s->pacc[3 * (size_t)p + a], SFX(ticks)(corr, s->pos_scale));
    }
    lam[1] = lt + (sticks ? tl / w : R(0.0));
  }
}

/* SDFCollider friction (SDFCollider.ts:271-291 + collider.ts:96-118) */
static void SFX(sdf_friction)(SFX(scene)* s, uint32_t p, REAL w, const REAL* xs, const REAL* v) {
  for (uint32_t k = 0; k < s->nsdf; ++k) {
    const SFX(sdfc)* D = &s->sdf[k];
    REAL ln = s->sdflam[((size_t)p * 4 + k) * 2];
    if (!(ln > R(0.0))) continue;
    REAL phi, g[3];
    SFX(sdf_sample)(D, xs, &phi, g);
    REAL gl = RSQRT(SFX(dot)(g, g));
    if (!(gl > (REAL)1e-6)) continue;
    REAL nrm[3] = {g[0] / gl, g[1] / gl, g[2] / gl}, vt[3];
    REAL nd = SFX(dot)(nrm, v);
    for (int a = 0; a < 3; ++a) vt[a] = v[a] - nrm[a] * nd;
    REAL vl = RSQRT(SFX(dot)(vt, vt));
    if (vl > (REAL)1e-6) {
      REAL change = RFMIN(((D->muk * ln) * w) / s->dt, vl);
      for (int a = 0; a < 3; ++a) {
        REAL d = -((vt[a] / vl) * change);
        s->vacc[3 * (size_t)p + a] = SFX(wadd)(s->vacc[3 * (size_t)p + a], SFX(ticks)(d, s->vel_scale));
      }
    }
  }
}

static void SFX(walls_solve)(SFX(scene)* s, REAL clock) {
  uint32_t n = s->n;
#pragma omp parallel for schedule(dynamic, 128)
  for (uint32_t p = 0; p < n; ++p) {
    REAL w = s->w[p];
    if (!(w > R(0.0))) continue;
    const REAL* xs = s->xs + 3 * (size_t)p;
    REAL disp[3] = {xs[0] - s->x[3 * (size_t)p], xs[1] - s->x[3 * (size_t)p + 1], xs[2] - s->x[3 * (size_t)p + 2]};
    for (uint32_t k = 0; k < s->nprim; ++k) {
      REAL phi, nrm[3];
      SFX(sdf)(&s->prim[k], xs, clock, &phi, nrm);
      if (phi < s->radius) {
        REAL cv[3], cd[3], slip[3], tang[3];
        SFX(surface_velocity)(&s->prim[k], xs, clock, cv);
        for (int a = 0; a < 3; ++a) cd[a] = cv[a] * s->dt;
        REAL depth = s->radius - phi;
        REAL* lam = s->plam + ((size_t)p * 32 + k) * 2;
        REAL ln = lam[0] + depth / w;
        lam[0] = ln;
        for (int a = 0; a < 3; ++a) slip[a] = disp[a] - cd[a];
        REAL sd = SFX(dot)(slip, nrm);
        for (int a = 0; a < 3; ++a) tang[a] = slip[a] - nrm[a] * sd;
        REAL tl = RSQRT(SFX(dot)(tang, tang));
        REAL lt = lam[1];
        int sticks = tl > (REAL)1e-10 && (tl / w) <= (s->prim[k].d1[3] * ln - lt);
        for (int a = 0; a < 3; ++a) {
          REAL push = nrm[a] * depth;
          REAL corr = sticks ? push - tang[a] : push;
          s->pacc[3 * (size_t)p + a] = SFX(wadd)(s->pacc[3 * (size_t)p + a], SFX(ticks)(corr, s->pos_scale));
        }
        lam[1] = lt + (sticks ? tl / w : R(0.0));
      }
    }
    if (s->nsdf) SFX(sdf_solve)(s, p, w, xs, disp);
  }
  int sat = 0;
#pragma omp parallel for schedule(static) reduction(|| : sat)
  for (uint32_t p = 0; p < n; ++p) {
    REAL c[3];
    if (SFX(acc_change)(s->pacc + 3 * (size_t)p, s->pos_inv, c)) sat = 1;
    for (int a = 0; a < 3; ++a) {
      s->xs[3 * (size_t)p + a] = s->xs[3 * (size_t)p + a] + c[a];
      s->pacc[3 * (size_t)p + a] = 0;
    }
  }
  if (sat) s->flag_pos = 1;
}

/* PrimitiveSet.ts:366-391 then the velocities accumulator apply (loop.ts:226-230) */
static void SFX(walls_friction)(SFX(scene)* s, REAL clock) {
  uint32_t n = s->n;
#pragma omp parallel for schedule(dynamic, 128)
  for (uint32_t p = 0; p < n; ++p) {
    REAL w = s->w[p];
    if (!(w > R(0.0))) continue;
    const REAL* xs = s->xs + 3 * (size_t)p;
    const REAL* v = s->v + 3 * (size_t)p;
    for (uint32_t k = 0; k < s->nprim; ++k) {
      REAL ln = s->plam[((size_t)p * 32 + k) * 2];
      if (!(ln > R(0.0))) continue;
      REAL phi, g[3];
      SFX(sdf)(&s->prim[k], xs, clock, &phi, g);
      REAL gl = RSQRT(SFX(dot)(g, g));
      if (!(gl > (REAL)1e-6)) continue;
      REAL nrm[3] = {g[0] / gl, g[1] / gl, g[2] / gl}, cv[3], rel[3], vt[3];
      SFX(surface_velocity)(&s->prim[k], xs, clock, cv);
      for (int a = 0; a < 3; ++a) rel[a] = v[a] - cv[a];
      REAL nd = SFX(dot)(nrm, rel);
      for (int a = 0; a < 3; ++a) vt[a] = rel[a] - nrm[a] * nd;
      REAL vl = RSQRT(SFX(dot)(vt, vt));
      if (vl > (REAL)1e-6) {
        REAL change = RFMIN(((s->prim[k].vel[3] * ln) * w) / s->dt, vl);
        for (int a = 0; a < 3; ++a) {
          REAL d = -((vt[a] / vl) * change);
          s->vacc[3 * (size_t)p + a] = SFX(wadd)(s->vacc[3 * (size_t)p + a], SFX(ticks)(d, s->vel_scale));
        }
      }
    }
    if (s->nsdf) SFX(sdf_friction)(s, p, w, xs, v);
  }
  int sat = 0;
#pragma omp parallel for schedule(static) reduction(|| : sat)
  for (uint32_t p = 0; p < n; ++p) {
    REAL c[3];
    if (SFX(acc_change)(s->vacc + 3 * (size_t)p, s->vel_inv, c)) sat = 1;
    for (int a = 0; a < 3; ++a) {
      s->v[3 * (size_t)p + a] = s->v[3 * (size_t)p + a] + c[a];
      s->vacc[3 * (size_t)p + a] = 0;
    }
  }
  if (sat) s->flag_vel = 1;
}

/* --------------------------------------------------------------------- */
/* global shape matching (shapeMatch.ts:53-351, polarDecomp.ts:34-264)   */
/* --------------------------------------------------------------------- */
static void SFX(polar)(const REAL a[9], REAL r[9]) {
#define A(i, j) a[3 * (i) + (j)]
  REAL m00 = (A(0, 0) * A(0, 0) + A(1, 0) * A(1, 0)) + A(2, 0) * A(2, 0);
  REAL m01 = (A(0, 0) * A(0, 1) + A(1, 0) * A(1, 1)) + A(2, 0) * A(2, 1);
  REAL m02 = (A(0, 0) * A(0, 2) + A(1, 0) * A(1, 2)) + A(2, 0) * A(2, 2);
  REAL m11 = (A(0, 1) * A(0, 1) + A(1, 1) * A(1, 1)) + A(2, 1) * A(2, 1);
  REAL m12 = (A(0, 1) * A(0, 2) + A(1, 1) * A(1, 2)) + A(2, 1) * A(2, 2);
  REAL m22 = (A(0, 2) * A(0, 2) + A(1, 2) * A(1, 2)) + A(2, 2) * A(2, 2);
  REAL q[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
  REAL eps = (REAL)1e-12;
  for (int sweep = 0; sweep < 8; ++sweep) {
    for (int plane = 0; plane < 3; ++plane) {
      int P = plane == 2 ? 1 : 0, Qi = plane == 0 ? 1 : 2;
      REAL* mpq = plane == 0 ? &m01 : (plane == 1 ? &m02 : &m12);
      REAL* mpp = P == 0 ? &m00 : &m11;
      REAL* mqq = Qi == 1 ? &m11 : &m22;
      if (!(RFABS(*mpq) > eps)) continue;
      REAL apq = *mpq, app = *mpp, aqq = *mqq;
      REAL tau = (aqq - app) / (apq * R(2.0));
      REAL sg = tau >= R(0.0) ? R(1.0) : R(-1.0);
      REAL ta = RFABS(tau);
      REAL t = sg / (ta + RSQRT(ta * ta + R(1.0)));
      REAL cc = R(1.0) / RSQRT(t * t + R(1.0));
      REAL ss = t * cc;
      /* the two other off-diagonals of the rotated plane */
      REAL *u, *v;
      if (plane == 0) { u = &m02; v = &m12; }
      else if (plane == 1) { u = &m01; v = &m12; }
      else { u = &m01; v = &m02; }
      REAL ou = *u, ov = *v;
      *mpp = app - t * apq;
      *mqq = aqq + t * apq;
      *mpq = R(0.0);
      *u = cc * ou - ss * ov;
      *v = ss * ou + cc * ov;
      for (int i = 0; i < 3; ++i) {
        REAL qp = q[i][P], qq = q[i][Qi];
        q[i][P] = cc * qp - ss * qq;
        q[i][Qi] = ss * qp + cc * qq;
      }
    }
  }
  REAL il[3] = {R(1.0) / RSQRT(RFMAX(m00, (REAL)1e-20)), R(1.0) / RSQRT(RFMAX(m11, (REAL)1e-20)),
                R(1.0) / RSQRT(RFMAX(m22, (REAL)1e-20))};
  REAL si[3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = i; j < 3; ++j) {
      REAL v = ((q[i][0] * q[j][0]) * il[0] + (q[i][1] * q[j][1]) * il[1]) + (q[i][2] * q[j][2]) * il[2];
      si[i][j] = v;
      si[j][i] = v;
    }
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) r[3 * i + j] = (A(i, 0) * si[0][j] + A(i, 1) * si[1][j]) + A(i, 2) * si[2][j];
#undef A
}

/* Tree fold of 256 per-thread partials, pairing i with i + s (shapeMatch.ts:88-101). */
static void SFX(tree)(REAL (*v)[13]) {
  for (int s = 128; s > 0; s >>= 1)
    for (int t = 0; t < s; ++t)
      for (int k = 0; k < 13; ++k) v[t][k] = v[t][k] + v[t + s][k];
}

static void SFX(shape_match)(SFX(scene)* s, int first, int pull) {
  static REAL part[256][13];
  for (uint32_t b = 0; b < s->nbodies; ++b) {
    uint32_t st = s->bstart[b], cnt = s->bcount[b];
    for (int t = 0; t < 256; ++t) {
      for (int k = 0; k < 13; ++k) part[t][k] = 0;
      for (uint32_t i = (uint32_t)t; i < cnt; i += 256) {
        uint32_t id = st + i;
        const REAL* x = s->xs + 3 * (size_t)id;
        REAL m = s->weights[id];
        const REAL* r = s->rest + 3 * (size_t)id;
        part[t][0] = part[t][0] + x[0] * m;
        part[t][1] = part[t][1] + x[1] * m;
        part[t][2] = part[t][2] + x[2] * m;
        part[t][3] = part[t][3] + m;
        REAL xm[3] = {x[0] * m, x[1] * m, x[2] * m};
        for (int row = 0; row < 3; ++row)
          for (int col = 0; col < 3; ++col) part[t][4 + 3 * row + col] = part[t][4 + 3 * row + col] + r[col] * xm[row];
      }
    }
    SFX(tree)(part);
    REAL c[3] = {part[0][0] / part[0][3], part[0][1] / part[0][3], part[0][2] / part[0][3]};
    REAL A[9], Rm[9];
    for (int k = 0; k < 9; ++k) A[k] = part[0][4 + k];
    SFX(polar)(A, Rm);
    for (int a = 0; a < 3; ++a) s->centers[b][a] = c[a];
    for (int k = 0; k < 9; ++k) s->rot[b][k] = Rm[k];
    if (!pull) continue;
    REAL at = s->bcomp[b] / (s->dt * s->dt);
    for (uint32_t i = 0; i < cnt; ++i) {
      uint32_t id = st + i;
      const REAL* r = s->rest + 3 * (size_t)id;
      REAL* x = s->xs + 3 * (size_t)id;
      REAL* lam = s->slam + 3 * (size_t)id;
      REAL w = s->w[id];
      REAL den = w + at;
      for (int a = 0; a < 3; ++a) {
        REAL goal = ((Rm[3 * a] * r[0] + Rm[3 * a + 1] * r[1]) + Rm[3 * a + 2] * r[2]) + c[a];
        REAL C = x[a] - goal;
        REAL lo = first ? R(0.0) : lam[a];
        REAL dl = ((-C) - lo * at) / den;
        lam[a] = lo + dl;
        x[a] = x[a] + dl * w;
      }
    }
  }
}

/* --------------------------------------------------------------------- */
/* local shape matching with oriented particles                          */
/* --------------------------------------------------------------------- */
/* buildNeighborGraph (SoftbodySystem.ts:470-495) and the rest centres of the
 * groups, accumulated into f32 storage (SoftbodySystem.ts:281-292). */
static void SFX(local_build)(SFX(scene)* s) {
  uint32_t n = s->n;
  uint32_t* deg = (uint32_t*)calloc(n, sizeof(uint32_t));
  s->span_lo = 0xffffffffu;
  s->span_hi = 0;
  for (uint32_t b = 0; b < s->nbodies; ++b) {
    for (uint32_t i = s->bstart[b]; i < s->bstart[b] + s->bcount[b]; ++i) deg[i] += 1;
    if (s->bstart[b] < s->span_lo) s->span_lo = s->bstart[b];
    if (s->bstart[b] + s->bcount[b] > s->span_hi) s->span_hi = s->bstart[b] + s->bcount[b];
  }
  for (uint32_t e = 0; e < s->nledges; ++e) {
    deg[s->ledges[2 * e]] += 1;
    deg[s->ledges[2 * e + 1]] += 1;
  }
  s->csr_off = (uint32_t*)calloc((size_t)n + 1, sizeof(uint32_t));
  for (uint32_t i = 0; i < n; ++i) s->csr_off[i + 1] = s->csr_off[i] + deg[i];
  uint32_t total = s->csr_off[n];
  s->csr_idx = (uint32_t*)calloc(total ? total : 1, sizeof(uint32_t));
  uint32_t* cur = deg;
  for (uint32_t i = 0; i < n; ++i) cur[i] = s->csr_off[i];
  /* per body: its particles, then its edges (edges are stored body by body) */
  uint32_t e = 0;
  for (uint32_t b = 0; b < s->nbodies; ++b) {
    uint32_t lo = s->bstart[b], hi = s->bstart[b] + s->bcount[b];
    for (uint32_t i = lo; i < hi; ++i) s->csr_idx[cur[i]++] = i;
    for (; e < s->nledges && s->ledges[2 * e] >= lo && s->ledges[2 * e] < hi; ++e) {
      uint32_t i = s->ledges[2 * e], j = s->ledges[2 * e + 1];
      s->csr_idx[cur[i]++] = j;
      s->csr_idx[cur[j]++] = i;
    }
  }
  free(deg);
  s->rcen = (REAL*)calloc(3 * (size_t)n, sizeof(REAL));
  s->gcen = (REAL*)calloc(3 * (size_t)n, sizeof(REAL));
  s->grot = (REAL*)calloc(9 * (size_t)n, sizeof(REAL));
  s->plam4 = (REAL*)calloc(3 * (size_t)(total ? total : 1), sizeof(REAL));
  s->lacc = (int32_t*)calloc(3 * (size_t)n, sizeof(int32_t));
  for (uint32_t b = 0; b < s->nbodies; ++b)
    for (uint32_t i = s->bstart[b]; i < s->bstart[b] + s->bcount[b]; ++i) {
      uint32_t from = s->csr_off[i], to = s->csr_off[i + 1];
      float acc[3] = {0.f, 0.f, 0.f};
      for (uint32_t k = from; k < to; ++k) {
        uint32_t j = s->csr_idx[k];
        for (int a = 0; a < 3; ++a) acc[a] = (float)((double)acc[a] + (double)(float)s->rest[3 * (size_t)j + a] / (double)(to - from));
      }
      for (int a = 0; a < 3; ++a) s->rcen[3 * (size_t)i + a] = (REAL)acc[a];
    }
}

/* rotation.ts:38-55 */
static void SFX(orient_predict)(SFX(scene)* s) {
  for (uint32_t i = s->span_lo; i < s->span_hi; ++i) {
    REAL* q = s->q + 4 * (size_t)i;
    REAL* qp = s->qp + 4 * (size_t)i;
    const REAL* om = s->om + 4 * (size_t)i;
    REAL mag = RSQRT(SFX(dot)(om, om));
    if (mag < (REAL)1e-8) {
      for (int a = 0; a < 4; ++a) qp[a] = q[a];
      continue;
    }
    REAL half = (mag * s->dt) * R(0.5);
    REAL sr = RSIN(half) / mag;
    REAL sp[3] = {om[0] * sr, om[1] * sr, om[2] * sr};
    REAL sw = RCOS(half), cr[3];
    SFX(cross)(sp, q, cr);
    for (int a = 0; a < 3; ++a) qp[a] = (q[a] * sw + sp[a] * q[3]) + cr[a];
    qp[3] = sw * q[3] - SFX(dot)(sp, q);
  }
}

/* shapeMatchImplicit.ts:47-402: centres, rotations, pulls, apply, orientations */
static void SFX(local_iteration)(SFX(scene)* s, int first) {
  const REAL ai = (REAL)(float)((s->radius_d * s->radius_d) / 5.0);
  const REAL maxc2 = (REAL)(float)((5.0 * s->radius_d) * (5.0 * s->radius_d));
  for (uint32_t i = s->span_lo; i < s->span_hi; ++i) {
    uint32_t o0 = s->csr_off[i], o1 = s->csr_off[i + 1], cnt = o1 - o0;
    REAL sum[3] = {0, 0, 0};
    for (uint32_t k = o0; k < o1; ++k)
      for (int a = 0; a < 3; ++a) sum[a] = sum[a] + s->xs[3 * (size_t)s->csr_idx[k] + a];
    REAL inv = R(1.0) / (REAL)(cnt > 1 ? cnt : 1);
    for (int a = 0; a < 3; ++a) s->gcen[3 * (size_t)i + a] = cnt ? sum[a] * inv : R(0.0);
  }
  for (uint32_t i = s->span_lo; i < s->span_hi; ++i) {
    uint32_t o0 = s->csr_off[i], o1 = s->csr_off[i + 1];
    REAL A[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, Rm[9];
    const REAL* ci = s->gcen + 3 * (size_t)i;
    const REAL* cb = s->rcen + 3 * (size_t)i;
    for (uint32_t k = o0; k < o1; ++k) {
      uint32_t j = s->csr_idx[k];
      const REAL* qj = s->qp + 4 * (size_t)j;
      REAL xx = qj[0] * qj[0], yy = qj[1] * qj[1], zz = qj[2] * qj[2];
      REAL xy = qj[0] * qj[1], xz = qj[0] * qj[2], yz = qj[1] * qj[2];
      REAL wx = qj[3] * qj[0], wy = qj[3] * qj[1], wz = qj[3] * qj[2];
      REAL rj[9] = {R(1.0) - (yy + zz) * R(2.0), (xy - wz) * R(2.0), (xz + wy) * R(2.0),
                    (xy + wz) * R(2.0), R(1.0) - (xx + zz) * R(2.0), (yz - wx) * R(2.0),
                    (xz - wy) * R(2.0), (yz + wx) * R(2.0), R(1.0) - (xx + yy) * R(2.0)};
      for (int t = 0; t < 9; ++t) A[t] = A[t] + ai * rj[t];
      REAL xv[3], tv[3];
      for (int a = 0; a < 3; ++a) {
        xv[a] = s->xs[3 * (size_t)j + a] - ci[a];
        tv[a] = s->rest[3 * (size_t)j + a] - cb[a];
      }
      for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) A[3 * r + c] = A[3 * r + c] + xv[r] * tv[c];
    }
    SFX(polar)(A, Rm);
    if (o1 == o0)
      for (int t = 0; t < 9; ++t) Rm[t] = (t % 4 == 0) ? R(1.0) : R(0.0);
    for (int t = 0; t < 9; ++t) s->grot[9 * (size_t)i + t] = Rm[t];
  }
  for (uint32_t i = s->span_lo; i < s->span_hi; ++i) {
    uint32_t o0 = s->csr_off[i], o1 = s->csr_off[i + 1];
    const REAL* Rm = s->grot + 9 * (size_t)i;
    const REAL* ci = s->gcen + 3 * (size_t)i;
    const REAL* cb = s->rcen + 3 * (size_t)i;
    REAL at = s->lcomp[i] / (s->dt * s->dt);
    for (uint32_t k = o0; k < o1; ++k) {
      uint32_t j = s->csr_idx[k];
      REAL rr[3], C[3], dl[3], dx[3];
      for (int a = 0; a < 3; ++a) rr[a] = s->rest[3 * (size_t)j + a] - cb[a];
      for (int a = 0; a < 3; ++a) {
        REAL goal = ((Rm[3 * a] * rr[0] + Rm[3 * a + 1] * rr[1]) + Rm[3 * a + 2] * rr[2]) + ci[a];
        C[a] = s->xs[3 * (size_t)j + a] - goal;
      }
      REAL wj = s->w[j];
      REAL den = wj + at;
      REAL* lam = s->plam4 + 3 * (size_t)k;
      for (int a = 0; a < 3; ++a) {
        REAL lo = first ? R(0.0) : lam[a];
        dl[a] = ((-C[a]) - lo * at) / den;
        lam[a] = lo + dl[a];
      }
      uint32_t mc = s->csr_off[j + 1] - s->csr_off[j];
      REAL mj = RFMAX((REAL)mc, R(1.0));
      for (int a = 0; a < 3; ++a) dx[a] = (dl[a] * wj) / mj;
      REAL sc = RSQRT(maxc2 / RFMAX(SFX(dot)(dx, dx), maxc2));
      for (int a = 0; a < 3; ++a)
        s->lacc[3 * (size_t)j + a] = SFX(wadd)(s->lacc[3 * (size_t)j + a], SFX(ticks)(dx[a] * sc, s->pos_scale));
    }
  }
  for (uint32_t j = s->span_lo; j < s->span_hi; ++j)
    for (int a = 0; a < 3; ++a) {
      s->xs[3 * (size_t)j + a] = s->xs[3 * (size_t)j + a] + (REAL)s->lacc[3 * (size_t)j + a] * s->pos_inv;
      s->lacc[3 * (size_t)j + a] = 0;
    }
  for (uint32_t i = s->span_lo; i < s->span_hi; ++i) {
    if (!(s->csr_off[i + 1] > s->csr_off[i])) continue;
    const REAL* M = s->grot + 9 * (size_t)i;
    REAL m00 = M[0], m01 = M[1], m02 = M[2], m10 = M[3], m11 = M[4], m12 = M[5], m20 = M[6], m21 = M[7], m22 = M[8];
    REAL trace = (m00 + m11) + m22, qx, qy, qz, qw;
    if (trace > R(0.0)) {
      REAL sq = R(0.5) / RSQRT(trace + R(1.0));
      qw = R(0.25) / sq; qx = (m21 - m12) * sq; qy = (m02 - m20) * sq; qz = (m10 - m01) * sq;
    } else if (m00 > m11 && m00 > m22) {
      REAL sq = R(2.0) * RSQRT(((R(1.0) + m00) - m11) - m22);
      qw = (m21 - m12) / sq; qx = R(0.25) * sq; qy = (m01 + m10) / sq; qz = (m02 + m20) / sq;
    } else if (m11 > m22) {
      REAL sq = R(2.0) * RSQRT(((R(1.0) + m11) - m00) - m22);
      qw = (m02 - m20) / sq; qx = (m01 + m10) / sq; qy = R(0.25) * sq; qz = (m12 + m21) / sq;
    } else {
      REAL sq = R(2.0) * RSQRT(((R(1.0) + m22) - m00) - m11);
      qw = (m10 - m01) / sq; qx = (m02 + m20) / sq; qy = (m12 + m21) / sq; qz = R(0.25) * sq;
    }
    const REAL* qs = s->q + 4 * (size_t)i;
    REAL d = ((qs[0] * qx + qs[1] * qy) + qs[2] * qz) + qs[3] * qw;
    REAL f = d < R(0.0) ? R(-1.0) : R(1.0);
    REAL* qp = s->qp + 4 * (size_t)i;
    qp[0] = qx * f; qp[1] = qy * f; qp[2] = qz * f; qp[3] = qw * f;
  }
}

/* rotation.ts:57-82 */
static void SFX(orient_advect)(SFX(scene)* s) {
  for (uint32_t i = s->span_lo; i < s->span_hi; ++i) {
    REAL* q = s->q + 4 * (size_t)i;
    const REAL* raw = s->qp + 4 * (size_t)i;
    REAL l = RSQRT(RFMAX(((raw[0] * raw[0] + raw[1] * raw[1]) + raw[2] * raw[2]) + raw[3] * raw[3], (REAL)1e-12));
    REAL qp[4] = {raw[0] / l, raw[1] / l, raw[2] / l, raw[3] / l};
    REAL qi[3] = {-q[0], -q[1], -q[2]}, cr[3], rv[3];
    SFX(cross)(qp, qi, cr);
    REAL rpre[3];
    for (int a = 0; a < 3; ++a) rpre[a] = (qp[a] * q[3] + qi[a] * qp[3]) + cr[a];
    REAL wpre = qp[3] * q[3] - SFX(dot)(qp, qi);
    REAL f = wpre < R(0.0) ? R(-1.0) : R(1.0);
    for (int a = 0; a < 3; ++a) rv[a] = rpre[a] * f;
    REAL rw = wpre * f;
    REAL hs = RSQRT(SFX(dot)(rv, rv));
    REAL* om = s->om + 4 * (size_t)i;
    if (hs < (REAL)1e-8) {
      om[0] = om[1] = om[2] = om[3] = R(0.0);
    } else {
      REAL angle = RATAN2(hs, rw) * R(2.0);
      REAL k = (angle / hs) / s->dt, o[3];
      for (int a = 0; a < 3; ++a) o[a] = rv[a] * k;
      REAL sc = RSQRT(R(10000.0) / RFMAX(SFX(dot)(o, o), R(10000.0)));
      for (int a = 0; a < 3; ++a) om[a] = o[a] * sc;
      om[3] = R(0.0);
    }
    for (int a = 0; a < 4; ++a) q[a] = qp[a];
  }
}

/* --------------------------------------------------------------------- */
/* particle contacts (src/core/contact, loop.ts:346-418)                 */
/* --------------------------------------------------------------------- */
/* LAMBDA_SCALE (ContactBuffer.ts:16): ticks per kg m of the multipliers */
#define ORC_CT_LSCALE 1048576.0

/* A particle searches for contacts unless a material keeps it quiet
 * (loop.ts:372-376): every fluid's range, plus the extra quiet ranges. */
static int SFX(ct_emitter)(const SFX(scene)* s, uint32_t i) {
  for (uint32_t f = 0; f < s->nfluids; ++f)
    if (i >= s->fluid[f].start && i - s->fluid[f].start < s->fluid[f].count) return 0;
  for (uint32_t q = 0; q < s->nquiet; ++q)
    if (i >= s->qstart[q] && i - s->qstart[q] < s->qcount[q]) return 0;
  return 1;
}

static int32_t SFX(ct_lticks)(REAL v) { return orc_f2i_sat((double)(v * (REAL)ORC_CT_LSCALE)); }

/* buildContactGenerateKernel (generate.ts:39-88): each emitter walks the 27
 * buckets around its prediction; a pair is kept once (by the emitter, the
 * lower index when both emit), never inside one non-zero collision group,
 * never between two pinned particles, and only within the candidate radius. */
static void SFX(contact_generate)(SFX(scene)* s) {
  uint32_t H = s->H;
  s->ct_count = 0;
  for (uint32_t i = 0; i < s->n; ++i) {
    if (!SFX(ct_emitter)(s, i)) continue;
    const REAL* xi = s->xs + 3 * (size_t)i;
    uint32_t gi = s->grp[i];
    REAL wi = s->w[i];
    int32_t q[3];
    for (int a = 0; a < 3; ++a) q[a] = orc_f2i_sat((double)RFLOOR((xi[a] - s->origin[a]) / s->cell));
    uint32_t visited[27], nv = 0;
    for (int dz = -1; dz <= 1; ++dz)
      for (int dy = -1; dy <= 1; ++dy)
        for (int dx = -1; dx <= 1; ++dx) {
          uint32_t b = orc_morton((int32_t)((uint32_t)q[0] + (uint32_t)dx), (int32_t)((uint32_t)q[1] + (uint32_t)dy),
                                  (int32_t)((uint32_t)q[2] + (uint32_t)dz)) & (H - 1);
          int seen = 0;
          for (uint32_t k = 0; k < nv; ++k) seen |= visited[k] == b;
          visited[nv++] = b;
          if (seen) continue;
          for (uint32_t t = s->start[b]; t < s->start[b + 1]; ++t) {
            uint32_t j = s->perm[t];
            if (j == i) continue;
            if (j < i && SFX(ct_emitter)(s, j)) continue;
            if (gi != 0u && gi == s->grp[j]) continue;
            if (wi <= R(0.0) && s->w[j] <= R(0.0)) continue;
            const REAL* xj = s->xs + 3 * (size_t)j;
            REAL d[3] = {xi[0] - xj[0], xi[1] - xj[1], xi[2] - xj[2]};
            if (SFX(dot)(d, d) >= s->ct_cr2) continue;
            uint32_t slot = s->ct_count++;
            if (slot < s->ct_max) {
              s->ct_i[slot] = i;
              s->ct_j[slot] = j;
            }
          }
        }
  }
  /* resetLambdaKernel (ContactBuffer.ts:67-71) */
  for (uint32_t c = 0; c < s->ct_max; ++c) {
    s->ct_ln[c] = 0;
    s->ct_lt[c] = 0;
  }
}

static uint32_t SFX(ct_used)(const SFX(scene)* s) { return s->ct_count < s->ct_max ? s->ct_count : s->ct_max; }

/* buildContactStabilizeKernel (solve.ts:134-181), then the positions
 * accumulator applied to the committed and the predicted positions
 * (loop.ts:389-393). */
static void SFX(contact_stabilize)(SFX(scene)* s) {
  uint32_t used = SFX(ct_used)(s);
  for (uint32_t c = 0; c < used; ++c) {
    uint32_t i = s->ct_i[c], j = s->ct_j[c];
    const REAL* xi = s->x + 3 * (size_t)i;
    const REAL* xj = s->x + 3 * (size_t)j;
    REAL off[3] = {xi[0] - xj[0], xi[1] - xj[1], xi[2] - xj[2]};
    REAL dist = RSQRT(SFX(dot)(off, off));
    REAL wi = s->w[i], wj = s->w[j], ws = wi + wj;
    if (dist <= (REAL)1e-8 || dist >= s->ct_dist || ws <= R(0.0)) continue;
    REAL nrm[3] = {off[0] / dist, off[1] / dist, off[2] / dist};
    REAL dl = (s->ct_dist - dist) / ws;
    for (int a = 0; a < 3; ++a) {
      s->pacc[3 * (size_t)i + a] = SFX(wadd)(s->pacc[3 * (size_t)i + a], SFX(ticks)(nrm[a] * (wi * dl), s->pos_scale));
      s->pacc[3 * (size_t)j + a] = SFX(wadd)(s->pacc[3 * (size_t)j + a], SFX(ticks)(-(nrm[a] * (wj * dl)), s->pos_scale));
    }
  }
  for (uint32_t p = 0; p < s->n; ++p) {
    REAL c[3];
    if (SFX(acc_change)(s->pacc + 3 * (size_t)p, s->pos_inv, c)) s->flag_pos = 1;
    for (int a = 0; a < 3; ++a) {
      s->x[3 * (size_t)p + a] = s->x[3 * (size_t)p + a] + c[a];
      s->xs[3 * (size_t)p + a] = s->xs[3 * (size_t)p + a] + c[a];
      s->pacc[3 * (size_t)p + a] = 0;
    }
  }
}

/* buildContactSolveKernel (solve.ts:33-96): one pass over the records; the
 * corrections go to the positions accumulator the colliders share. */
static void SFX(contact_solve)(SFX(scene)* s) {
  uint32_t used = SFX(ct_used)(s);
  for (uint32_t c = 0; c < used; ++c) {
    uint32_t i = s->ct_i[c], j = s->ct_j[c];
    const REAL* pi = s->xs + 3 * (size_t)i;
    const REAL* pj = s->xs + 3 * (size_t)j;
    REAL off[3] = {pi[0] - pj[0], pi[1] - pj[1], pi[2] - pj[2]};
    REAL dist = RSQRT(SFX(dot)(off, off));
    REAL wi = s->w[i], wj = s->w[j], ws = wi + wj;
    if (dist <= (REAL)1e-8 || dist >= s->ct_dist || ws <= R(0.0)) continue;
    REAL nrm[3] = {off[0] / dist, off[1] / dist, off[2] / dist};
    REAL depth = s->ct_dist - dist;
    REAL dln = depth / ws;
    s->ct_ln[c] = SFX(wadd)(s->ct_ln[c], SFX(ct_lticks)(dln));
    for (int a = 0; a < 3; ++a) s->ct_n[3 * (size_t)c + a] = nrm[a];
    const REAL* ci = s->x + 3 * (size_t)i;
    const REAL* cj = s->x + 3 * (size_t)j;
    REAL slip[3], tang[3];
    for (int a = 0; a < 3; ++a) slip[a] = (pi[a] - ci[a]) - (pj[a] - cj[a]);
    REAL sd = SFX(dot)(slip, nrm);
    for (int a = 0; a < 3; ++a) tang[a] = slip[a] - nrm[a] * sd;
    REAL tl = RSQRT(SFX(dot)(tang, tang));
    REAL ln = (REAL)s->ct_ln[c] / (REAL)ORC_CT_LSCALE;
    REAL lt = (REAL)s->ct_lt[c] / (REAL)ORC_CT_LSCALE;
    REAL dlt = tl / ws;
    int sticks = tl > (REAL)1e-10 && dlt <= s->ct_mus * ln - lt;
    for (int a = 0; a < 3; ++a) {
      REAL ni = nrm[a] * (wi * dln);
      REAL nj = -(nrm[a] * (wj * dln));
      REAL fi = -(tang[a] * (wi / ws));
      REAL fj = tang[a] * (wj / ws);
      s->pacc[3 * (size_t)i + a] = SFX(wadd)(s->pacc[3 * (size_t)i + a], SFX(ticks)(sticks ? ni + fi : ni, s->pos_scale));
      s->pacc[3 * (size_t)j + a] = SFX(wadd)(s->pacc[3 * (size_t)j + a], SFX(ticks)(sticks ? nj + fj : nj, s->pos_scale));
    }
    s->ct_lt[c] = SFX(wadd)(s->ct_lt[c], SFX(ct_lticks)(sticks ? dlt : R(0.0)));
  }
}

/* buildContactFrictionKernel (solve.ts:183-189 and 147-182): kinetic
 * friction into the velocities accumulator. */
static void SFX(contact_friction)(SFX(scene)* s) {
  uint32_t used = SFX(ct_used)(s);
  for (uint32_t c = 0; c < used; ++c) {
    uint32_t i = s->ct_i[c], j = s->ct_j[c];
    REAL wi = s->w[i], wj = s->w[j], ws = wi + wj;
    REAL ln = (REAL)s->ct_ln[c] / (REAL)ORC_CT_LSCALE;
    if (ws <= R(0.0) || ln <= R(0.0)) continue;
    const REAL* nrm = s->ct_n + 3 * (size_t)c;
    const REAL* vi = s->v + 3 * (size_t)i;
    const REAL* vj = s->v + 3 * (size_t)j;
    REAL rel[3] = {vi[0] - vj[0], vi[1] - vj[1], vi[2] - vj[2]}, vt[3];
    REAL nd = SFX(dot)(nrm, rel);
    for (int a = 0; a < 3; ++a) vt[a] = rel[a] - nrm[a] * nd;
    REAL vl = RSQRT(SFX(dot)(vt, vt));
    if (vl < (REAL)1e-6) continue;
    REAL change = RFMIN(((s->ct_muk * ln) * ws) / s->dt, vl);
    for (int a = 0; a < 3; ++a) {
      REAL imp = (-((vt[a] / vl) * change)) / ws;
      s->vacc[3 * (size_t)i + a] = SFX(wadd)(s->vacc[3 * (size_t)i + a], SFX(ticks)(imp * wi, s->vel_scale));
      s->vacc[3 * (size_t)j + a] = SFX(wadd)(s->vacc[3 * (size_t)j + a], SFX(ticks)(-(imp * wj), s->vel_scale));
    }
  }
}

/* --------------------------------------------------------------------- */
/* surface field (src/fluids/render/SurfaceField.ts:101-369)             */
/* --------------------------------------------------------------------- */
/* FIXED (SurfaceField.ts:32) and FIELD_BAND (34) */
#define ORC_FIELD_FIXED 65536.0
#define ORC_FIELD_BAND 6.0

/* The field's constants, as the reference turns its JS doubles into the
 * f32 literals and uniforms the kernels read. */
typedef struct {
  uint32_t n[3], count, taps, samples;
  uint32_t fluid_start, fluid_count, solid_start, solid_count, wet;
  REAL origin[3], dims[3], voxel, inv_voxel;
  REAL weights[13];
  REAL stretch, reach3, inv_samples, vol_scale, solid_scale;
  REAL radius, men_h, men_w;
} SFX(fieldp);

static inline uint32_t SFX(fidx)(const SFX(fieldp)* F, REAL cx, REAL cy, REAL cz) {
  /* indexOf over float cell coordinates (SurfaceField.ts:147-152) */
  REAL v = (cz * (REAL)(F->n[0] * F->n[1]) + cy * (REAL)F->n[0]) + cx;
  return (uint32_t)v;
}

/* splat (SurfaceField.ts:174-214): cloud-in-cell weights of each sample in
 * fixed point; `fluid` adds mass and mass-weighted offsets, else solid mass. */
static void SFX(field_splat)(const SFX(scene)* s, const SFX(fieldp)* F, int32_t* mass, int32_t* off, int32_t* solid,
                             uint32_t start, uint32_t count, int fluid) {
  for (uint32_t i = start; i < start + count; ++i) {
    const REAL* pos = s->x + 3 * (size_t)i;
    uint32_t samples = fluid ? F->samples : 1u;
    REAL trail[3] = {0, 0, 0};
    if (samples > 1) {
      REAL vv[3];
      for (int a = 0; a < 3; ++a) vv[a] = s->v[3 * (size_t)i + a] * F->stretch;
      REAL len = RSQRT(SFX(dot)(vv, vv));
      REAL k = F->reach3 / RFMAX(len, F->reach3);
      for (int a = 0; a < 3; ++a) trail[a] = vv[a] * k;
    }
    for (uint32_t sm = 0; sm < samples; ++sm) {
      REAL at[3], g[3];
      REAL frac = samples > 1 ? (REAL)((double)sm / (double)(samples - 1)) : R(0.0);
      for (int a = 0; a < 3; ++a) at[a] = sm == 0 ? pos[a] : pos[a] - trail[a] * frac;
      for (int a = 0; a < 3; ++a) g[a] = (at[a] - F->origin[a]) * F->inv_voxel - R(0.5);
      int inside = 1;
      for (int a = 0; a < 3; ++a) inside &= g[a] >= R(0.0) && g[a] < F->dims[a] - R(1.0);
      if (!inside) continue;
      REAL b[3], f[3];
      for (int a = 0; a < 3; ++a) { b[a] = RFLOOR(g[a]); f[a] = g[a] - b[a]; }
      for (int z = 0; z < 2; ++z)
        for (int y = 0; y < 2; ++y)
          for (int x = 0; x < 2; ++x) {
            REAL w = (((x ? f[0] : R(1.0) - f[0]) * (y ? f[1] : R(1.0) - f[1])) * (z ? f[2] : R(1.0) - f[2])) *
                     (fluid ? F->inv_samples : R(1.0));
            uint32_t c = SFX(fidx)(F, b[0] + (REAL)x, b[1] + (REAL)y, b[2] + (REAL)z);
            if (fluid) {
              mass[c] = SFX(wadd)(mass[c], orc_f2i_sat((double)RRINT(w * (REAL)ORC_FIELD_FIXED)));
              off[3 * (size_t)c + 0] = SFX(wadd)(off[3 * (size_t)c + 0], orc_f2i_sat((double)RRINT((w * (f[0] - (REAL)x)) * (REAL)ORC_FIELD_FIXED)));
              off[3 * (size_t)c + 1] = SFX(wadd)(off[3 * (size_t)c + 1], orc_f2i_sat((double)RRINT((w * (f[1] - (REAL)y)) * (REAL)ORC_FIELD_FIXED)));
              off[3 * (size_t)c + 2] = SFX(wadd)(off[3 * (size_t)c + 2], orc_f2i_sat((double)RRINT((w * (f[2] - (REAL)z)) * (REAL)ORC_FIELD_FIXED)));
            } else {
              solid[c] = SFX(wadd)(solid[c], orc_f2i_sat((double)RRINT(w * (REAL)ORC_FIELD_FIXED)));
            }
          }
    }
  }
}

/* wallKernel (SurfaceField.ts:251-285): wetting proximity and the nearest
 * collider distance per voxel, from the scene's primitives and SDF colliders. */
static void SFX(field_walls)(const SFX(scene)* s, const SFX(fieldp)* F, REAL* walls) {
  for (uint32_t i = 0; i < F->count; ++i) {
    REAL prox = R(0.0), nearest = R(1e3);
    if (F->wet && (s->nprim || s->nsdf)) {
      uint32_t c[3] = {i % F->n[0], (i / F->n[0]) % F->n[1], i / (F->n[0] * F->n[1])};
      REAL p[3];
      for (int a = 0; a < 3; ++a) p[a] = F->origin[a] + ((REAL)c[a] + R(0.5)) * F->voxel;
      for (uint32_t k = 0; k < s->nprim; ++k) {
        REAL phi, g[3];
        SFX(prim_shape)(&s->prim[k], p, &phi, g);
        nearest = RFMIN(nearest, phi);
      }
      for (uint32_t k = 0; k < s->nsdf; ++k) {
        const SFX(sdfc)* D = &s->sdf[k];
        REAL phi, g[3];
        SFX(sdf_sample)(D, p, &phi, g);
        REAL half[3], bo[3], ctr[3], d[3];
        for (int a = 0; a < 3; ++a) { half[a] = (D->res[a] * D->voxel[a]) * R(0.5); bo[a] = D->origin[a] + half[a]; }
        for (int r = 0; r < 3; ++r)
          ctr[r] = D->pos[r] + ((D->rot[3 * r] * bo[0] + D->rot[3 * r + 1] * bo[1]) + D->rot[3 * r + 2] * bo[2]) * D->scale;
        for (int a = 0; a < 3; ++a) d[a] = p[a] - ctr[a];
        REAL bound = RSQRT(SFX(dot)(d, d)) - RSQRT(SFX(dot)(half, half)) * D->scale;
        nearest = RFMIN(nearest, RFMAX(phi, bound));
      }
      REAL width = RFMAX(F->men_w, (REAL)1e-5);
      prox = REXP(-(RFMAX(nearest, R(0.0)) / width));
    }
    walls[2 * (size_t)i] = prox;
    walls[2 * (size_t)i + 1] = nearest;
  }
}

/* The three separable Gaussian passes and the final field
 * (SurfaceField.ts:286-353); out: (phi in voxels, volume fraction,
 * wall proximity, 1) per voxel. */
static void SFX(field_blur)(const SFX(fieldp)* F, const int32_t* mass, const int32_t* off, const int32_t* solid,
                            const REAL* walls, REAL* out) {
  uint32_t count = F->count, nx = F->n[0], ny = F->n[1];
  REAL* a = (REAL*)calloc(4 * (size_t)count, sizeof(REAL));
  REAL* b = (REAL*)calloc(4 * (size_t)count, sizeof(REAL));
  REAL* sa = (REAL*)calloc(count, sizeof(REAL));
  REAL* sb = (REAL*)calloc(count, sizeof(REAL));
  for (int axis = 0; axis < 3; ++axis) {
    for (uint32_t i = 0; i < count; ++i) {
      int32_t c[3] = {(int32_t)(i % nx), (int32_t)((i / nx) % ny), (int32_t)(i / (nx * ny))};
      REAL sum[4] = {0, 0, 0, 0}, ssum = R(0.0);
      for (int k = -(int)F->taps; k <= (int)F->taps; ++k) {
        int32_t nb[3] = {c[0], c[1], c[2]};
        nb[axis] += k;
        if (nb[axis] < 0 || nb[axis] >= (int32_t)F->n[axis]) continue;
        uint32_t j = (uint32_t)nb[2] * (nx * ny) + (uint32_t)nb[1] * nx + (uint32_t)nb[0];
        REAL sm[4], ss;
        if (axis == 0) {
          sm[0] = (REAL)mass[j] / (REAL)ORC_FIELD_FIXED;
          for (int q = 0; q < 3; ++q) sm[1 + q] = (REAL)off[3 * (size_t)j + q] / (REAL)ORC_FIELD_FIXED;
          ss = (REAL)solid[j] / (REAL)ORC_FIELD_FIXED;
        } else {
          const REAL* src = axis == 1 ? a : b;
          for (int q = 0; q < 4; ++q) sm[q] = src[4 * (size_t)j + q];
          ss = (axis == 1 ? sa : sb)[j];
        }
        REAL reb[4] = {sm[0], sm[1], sm[2], sm[3]};
        for (int q = 0; q < 3; ++q) reb[1 + q] = sm[1 + q] + (REAL)(q == axis ? k : 0) * sm[0];
        REAL wk = F->weights[k + (int)F->taps];
        for (int q = 0; q < 4; ++q) sum[q] = sum[q] + reb[q] * wk;
        ssum = ssum + ss * wk;
      }
      if (axis < 2) {
        REAL* dst = axis == 0 ? a : b;
        for (int q = 0; q < 4; ++q) dst[4 * (size_t)i + q] = sum[q];
        (axis == 0 ? sa : sb)[i] = ssum;
        continue;
      }
      REAL w = sum[0];
      REAL den = RFMAX(w, (REAL)1e-8);
      REAL mean[3] = {sum[1] / den, sum[2] / den, sum[3] / den};
      REAL radius = F->radius * F->inv_voxel;
      REAL phi = w > (REAL)1e-5 ? RSQRT(SFX(dot)(mean, mean)) - radius : (REAL)ORC_FIELD_BAND;
      phi = RFMIN(phi, (REAL)ORC_FIELD_BAND);
      REAL sp = RFMIN(RFMAX(ssum * F->solid_scale, R(0.0)), R(1.0));
      REAL prox = RFMAX(sp, walls[2 * (size_t)i]);
      REAL lift = F->men_h * F->inv_voxel;
      REAL reach = RFMIN(RFMAX(R(1.0) - phi / RFMAX(lift * R(1.5), (REAL)1e-5), R(0.0)), R(1.0));
      phi = phi - (lift * prox) * (reach * reach);
      phi = RFMAX(phi, -walls[2 * (size_t)i + 1] * F->inv_voxel);
      out[4 * (size_t)i + 0] = phi;
      out[4 * (size_t)i + 1] = RFMIN(w * F->vol_scale, R(4.0));
      out[4 * (size_t)i + 2] = prox;
      out[4 * (size_t)i + 3] = R(1.0);
    }
  }
  free(a); free(b); free(sa); free(sb);
}

static void SFX(surface_field)(const SFX(scene)* s, const SFX(fieldp)* F, REAL* out) {
  int32_t* mass = (int32_t*)calloc(F->count, sizeof(int32_t));
  int32_t* off = (int32_t*)calloc(3 * (size_t)F->count, sizeof(int32_t));
  int32_t* solid = (int32_t*)calloc(F->count, sizeof(int32_t));
  REAL* walls = (REAL*)calloc(2 * (size_t)F->count, sizeof(REAL));
  SFX(field_splat)(s, F, mass, off, solid, F->fluid_start, F->fluid_count, 1);
  if (F->solid_count) SFX(field_splat)(s, F, mass, off, solid, F->solid_start, F->solid_count, 0);
  SFX(field_walls)(s, F, walls);
  SFX(field_blur)(F, mass, off, solid, walls, out);
  free(mass); free(off); free(solid); free(walls);
}

/* --------------------------------------------------------------------- */
/* one step (loop.ts:175-279)                                            */
/* --------------------------------------------------------------------- */
static void SFX(step)(SFX(scene)* s, double dt_step) {
  uint32_t n = s->n;
  s->dt = (REAL)(dt_step / (double)s->substeps);
  {
    REAL clock = s->dt * ((REAL)s->substeps - R(1.0));
    for (uint32_t k = 0; k < s->substeps; ++k) {
      s->clock[k] = clock;
      clock = RFMAX(clock - s->dt, R(0.0));
    }
  }
  s->flag_pos = 0;
  s->flag_vel = 0;
  if (s->local && !s->csr_off) SFX(local_build)(s);
  if (!s->initialized) {
    /* first step: static boundary volumes from committed positions, over a
     * grid of the current predictions (loop.ts:236, 276-277) */
    int any_static = 0;
    for (uint32_t f = 0; f < s->nfluids; ++f)
      for (uint32_t b = 0; b < s->fluid[f].nb; ++b) any_static |= !s->fluid[f].bdyn[b];
    if (any_static) {
      for (uint32_t i = 0; i < n; ++i)
        if (s->w[i] > R(0.0))
          for (int a = 0; a < 3; ++a) s->xs[3 * (size_t)i + a] = s->x[3 * (size_t)i + a];
      SFX(grid_build)(s);
      for (uint32_t f = 0; f < s->nfluids; ++f)
        for (uint32_t b = 0; b < s->fluid[f].nb; ++b)
          if (!s->fluid[f].bdyn[b]) SFX(boundary_volume)(s, &s->fluid[f], b, s->x);
    }
    s->initialized = 1;
  }
  for (uint32_t k = 0; k < s->substeps; ++k) {
    /* predict (integrate.ts:25-34) */
#pragma omp parallel for schedule(static)
    for (uint32_t i = 0; i < n; ++i) {
      if (!(s->w[i] > R(0.0))) continue;
      for (int a = 0; a < 3; ++a) {
        REAL vel = s->v[3 * (size_t)i + a] + s->g[a] * s->dt;
        s->v[3 * (size_t)i + a] = vel;
        s->xs[3 * (size_t)i + a] = s->x[3 * (size_t)i + a] + vel * s->dt;
      }
    }
    if (s->local) SFX(orient_predict)(s);
    s->flag_nl = 0;
    if (s->nfluids || s->contact) SFX(grid_build)(s);
    if (s->contact) {
      SFX(contact_generate)(s);
      SFX(contact_stabilize)(s);
    }
    for (uint32_t f = 0; f < s->nfluids; ++f)
      for (uint32_t b = 0; b < s->fluid[f].nb; ++b)
        if (s->fluid[f].bdyn[b]) SFX(boundary_volume)(s, &s->fluid[f], b, s->xs);
    memset(s->plam, 0, sizeof(REAL) * 64 * (size_t)n);
    if (s->sdflam) memset(s->sdflam, 0, sizeof(REAL) * 8 * (size_t)n);
    for (uint32_t f = 0; f < s->nfluids; ++f) SFX(neighbors)(s, &s->fluid[f]);
    for (uint32_t f = 0; f < s->nfluids; ++f) {
      SFX(fluid)* F = &s->fluid[f];
      if (F->effects & 4u) SFX(color_normal)(s, F);
      if (F->effects & (4u | 8u)) SFX(impulses)(s, F);
    }
    for (uint32_t it = 0; it < s->iterations; ++it) {
      for (uint32_t f = 0; f < s->nfluids; ++f) {
        SFX(lambda_pass)(s, &s->fluid[f]);
        SFX(delta_pass)(s, &s->fluid[f]);
      }
      if (s->local) SFX(local_iteration)(s, it == 0);
      else if (s->nbodies) SFX(shape_match)(s, it == 0, 1);
      if (s->contact) SFX(contact_solve)(s);
      if (s->nprim || s->nsdf || s->contact) SFX(walls_solve)(s, s->clock[k]);
    }
    /* advect (integrate.ts:36-44) */
#pragma omp parallel for schedule(static)
    for (uint32_t i = 0; i < n; ++i) {
      if (!(s->w[i] > R(0.0))) continue;
      for (int a = 0; a < 3; ++a) {
        s->v[3 * (size_t)i + a] = (s->xs[3 * (size_t)i + a] - s->x[3 * (size_t)i + a]) / s->dt;
        s->x[3 * (size_t)i + a] = s->xs[3 * (size_t)i + a];
      }
    }
    for (uint32_t f = 0; f < s->nfluids; ++f) {
      SFX(velocity_passes)(s, &s->fluid[f]);
      if (s->fluid[f].isv_iters) SFX(viscosity_solve)(s, &s->fluid[f]);
    }
    if (s->local) {
      SFX(orient_advect)(s);
      SFX(shape_match)(s, 0, 0);
    }
    if ((s->nprim || s->nsdf || s->contact) && s->iterations > 0) {
      if (s->contact) SFX(contact_friction)(s);
      SFX(walls_friction)(s, s->clock[k]);
    }
  }
}
