// FluidSystem, PrimitiveSet, SoftbodySystem, SimLoop and the batched form
// behind the C ABI, plus the substep schedule shared by loops and batches,
// and the Simulation scene-layout helpers.
#include <algorithm>
#include <cmath>

#include "runtime.hpp"

namespace pf {
thread_local uint32_t tile_grid_cap = 0;

void particles_scatter(pf_particles* ps, uint32_t start, uint32_t count, const std::vector<float4>* pos,
                       const std::vector<float4>* vel, const std::vector<float>* f,
                       const std::vector<uint32_t>* grp, uint32_t mask);
void particles_gather(pf_particles* ps, int what, void* host, size_t elem);

// three's DataUtils.toHalfFloat (three@0.184.0): clamp to +-65504, then the
// base/shift tables of van der Zijp's fast half-float conversion, which
// truncate the mantissa.
static uint16_t half_from_float(float v) {
  static uint16_t base[512];
  static uint8_t shift[512];
  static bool ready = false;
  if (!ready) {
    for (int i = 0; i < 256; ++i) {
      const int e = i - 127;
      uint16_t b;
      uint8_t sh;
      if (e < -27) { b = 0x0000; sh = 24; }
      else if (e < -14) { b = (uint16_t)(0x0400 >> (-e - 14)); sh = (uint8_t)(-e - 1); }
      else if (e <= 15) { b = (uint16_t)((e + 15) << 10); sh = 13; }
      else if (e < 128) { b = 0x7c00; sh = 24; }
      else { b = 0x7c00; sh = 13; }
      base[i] = b;
      base[i | 0x100] = (uint16_t)(b | 0x8000);
      shift[i] = shift[i | 0x100] = sh;
    }
    ready = true;
  }
  if (v > 65504.0f) v = 65504.0f;
  if (v < -65504.0f) v = -65504.0f;
  uint32_t f;
  std::memcpy(&f, &v, 4);
  const uint32_t e = (f >> 23) & 0x1ffu;
  return (uint16_t)(base[e] + ((f & 0x007fffffu) >> shift[e]));
}

// Predict without a grid (integrate.ts:25-34), for scenes with no neighbour
// queries: the state keeps its order and the prediction goes to xs[c ^ 1].
template <bool PERSIST>
__global__ void k_predict(const SceneDev* __restrict__ scenes, const Tile* __restrict__ tiles,
                          uint32_t ntiles, int p, int c, int reset_step_flags) {
  pf_for_tiles<PERSIST>(ntiles, [&](uint32_t ti) {
    const Tile t = tiles[ti];
    const SceneDev& S = scenes[t.scene];
    const uint32_t s = t.first + threadIdx.x;
    if (reset_step_flags && t.first == 0 && threadIdx.x == 0) {
      S.flags[PF_FLAG_POSITIONS] = 0u;
      S.flags[PF_FLAG_VELOCITIES] = 0u;
    }
    if (s >= S.n) return;
    const float w = S.w[p][s];
    float4 xs = S.xs[c][s];
    if (w > 0.0f) {
      const float4 v = S.v[p][s];
      const float vx = v.x + S.gx * S.dt, vy = v.y + S.gy * S.dt, vz = v.z + S.gz * S.dt;
      S.v[p][s] = make_float4(vx, vy, vz, 0.0f);
      const float4 x = S.x[p][s];
      xs = make_float4(x.x + vx * S.dt, x.y + vy * S.dt, x.z + vz * S.dt, 0.0f);
    }
    S.xs[c ^ 1][s] = make_float4(xs.x, xs.y, xs.z, 0.0f);
  });
}

// The substep schedule (loop.ts:175-241, FluidSystem.ts:200-309), fused as
// described in fluid.hip. Scenes start every step at state parity 0 and
// prediction buffer 0; the caller swaps descriptor variants afterwards.
// Returns the number of kernel launches.
uint32_t enqueue_step(Context* ctx, const SceneDev* scenes, const PlanShape& sh, const Tile* tiles,
                      uint32_t ntiles, const uint2* stiles, uint32_t nstiles, const uint2* btiles,
                      uint32_t nbtiles, const uint2* ltiles, uint32_t nltiles, const Tile* ctiles,
                      uint32_t nctiles, void* scan_ctl) {
  hipStream_t st = ctx->stream;
  const TileGridScope launch_shape(sh.grid_cap);
  uint32_t nk = 0;
  int c = 0;
  const bool grid = sh.grid;
  enum LastOp { NONE, DELTA, WALLS, SOFT };
  // The last fluid's correction pass can carry the container contact and the
  // advect only when nothing runs after it inside an iteration (soft bodies,
  // a boundary reaction apply).
  bool reaction = false;
  for (uint32_t f = 0; f < sh.nfluids; ++f) reaction = reaction || sh.nb[f] > 0;
  // Contacts solve after every material and before the colliders, so the
  // accumulator apply cannot ride on the fluid's correction pass.
  const bool delta_can_end = !sh.soft && !reaction && sh.nfluids > 0 && !sh.contact;
  const bool apply = sh.walls || sh.contact;  // a positions/velocities accumulator exists
  LastOp last = NONE;
  if (apply) last = delta_can_end ? DELTA : WALLS;
  else if (sh.soft) last = SOFT;
  else if (delta_can_end) last = DELTA;
  for (uint32_t k = 0; k < sh.substeps; ++k) {
    int p = 0;
    if (grid) {
      p = (int)(k & 1u);
      { PF_LAUNCH(ctx, "hash_predict"); launch_hash(st, scenes, tiles, ntiles, p, c, k == 0); ++nk; }
      { PF_LAUNCH(ctx, "scan"); launch_scan(st, scenes, stiles, nstiles, scan_ctl); ++nk; }
      { PF_LAUNCH(ctx, "place"); launch_place(st, scenes, tiles, ntiles, p); ++nk; }
      { PF_LAUNCH(ctx, "rank_permute"); launch_rank(st, scenes, tiles, ntiles, p, c, c ^ 1); ++nk; }
      c ^= 1;
      p ^= 1;
      if (sh.contact) {
        // Contacts (loop.ts:370-393): the pair search on this substep's
        // predictions, then the stabilisation, whose predictions go to the
        // other buffer so the grid's sorted copy stays for the neighbour build.
        { PF_LAUNCH(ctx, "contact_generate"); launch_contact_generate(st, scenes, tiles, ntiles, p, c); ++nk; }
        {
          PF_LAUNCH(ctx, "contact_stabilize");
          launch_contact_stabilize(st, scenes, tiles, ntiles, ctiles, nctiles, p, c);
          nk += 2;
        }
        c ^= 1;
      }
    } else {
      PF_LAUNCH(ctx, "predict");
      launch_tiles(k_predict<false>, k_predict<true>, ntiles, 0, st, scenes, tiles, ntiles, 0, c, k == 0);
      ++nk;
      c ^= 1;
    }
    if (sh.local_soft) {
      PF_LAUNCH(ctx, "orient_predict");
      launch_local_predict(st, scenes, ltiles, nltiles);
      ++nk;
    }
    for (uint32_t f = 0; f < sh.nfluids; ++f) {
      if (sh.bdyn[f]) {
        PF_LAUNCH(ctx, "boundary_volume");
        launch_boundary_volume(st, scenes, tiles, ntiles, (int)f, p, c, 1);
        ++nk;
      }
    }
    for (uint32_t f = 0; f < sh.nfluids; ++f) {
      PF_LAUNCH(ctx, "neighbors");
      launch_neighbors(st, scenes, tiles, ntiles, (int)f, p, c, (sh.fx[f] & PF_FX_SURFACE_TENSION) != 0,
                       sh.tu_cap[f], sh.contact ? 1 : 0);
      ++nk;
    }
    for (uint32_t f = 0; f < sh.nfluids; ++f) {
      const bool stx = (sh.fx[f] & PF_FX_SURFACE_TENSION) != 0;
      const bool adh = (sh.fx[f] & PF_FX_ADHESION) != 0;
      if (stx || adh) {
        {
          PF_LAUNCH(ctx, "impulses");
          launch_impulses(st, scenes, tiles, ntiles, (int)f, p, c, stx, adh && sh.any_boundary);
          ++nk;
        }
        if (stx) {
          PF_LAUNCH(ctx, "impulse_apply");
          launch_impulse_apply(st, scenes, tiles, ntiles, (int)f, p, c);
          ++nk;
        }
        c ^= 1;
      }
    }
    for (uint32_t it = 0; it < sh.iterations; ++it) {
      const bool lastit = it + 1 == sh.iterations;
      for (uint32_t f = 0; f < sh.nfluids; ++f) {
        const bool lastf = f + 1 == sh.nfluids;
        { PF_LAUNCH(ctx, "lambda"); launch_lambda(st, scenes, tiles, ntiles, (int)f, p, c, sh.tu_cap[f]); ++nk; }
        const bool fuse_walls = lastf && last == DELTA && sh.walls;
        const bool fuse_adv = lastf && lastit && last == DELTA;
        {
          PF_LAUNCH(ctx, fuse_walls ? "delta_walls" : "delta");
          launch_delta(st, scenes, tiles, ntiles, (int)f, p, c, fuse_walls, fuse_adv, it == 0, (int)k,
                       sh.tu_cap[f]);
          ++nk;
        }
        c ^= 1;
        if (sh.nb[f]) {
          PF_LAUNCH(ctx, "reaction_apply");
          launch_reaction_apply(st, scenes, tiles, ntiles, (int)f, p, c);
          ++nk;
        }
      }
      if (sh.soft) {
        PF_LAUNCH(ctx, "shape_match");
        launch_shape_match(st, scenes, btiles, nbtiles, p, c, it == 0, 1);
        ++nk;
      }
      if (sh.local_soft) {
        PF_LAUNCH(ctx, "local_shape_match");
        launch_local_solve(st, scenes, ltiles, nltiles, p, c, it == 0);
        nk += 5;
      }
      if (sh.contact) {
        PF_LAUNCH(ctx, "contact_solve");
        launch_contact_solve(st, scenes, ctiles, nctiles, p, c);
        ++nk;
      }
      if (apply && last == WALLS) {
        PF_LAUNCH(ctx, "walls");
        launch_walls(st, scenes, tiles, ntiles, p, c, lastit, it == 0, (int)k);
        ++nk;
      }
    }
    if (sh.iterations == 0 || last == SOFT || last == NONE) {
      PF_LAUNCH(ctx, "advect");
      launch_advect(st, scenes, tiles, ntiles, p, c);
      ++nk;
    }
    if (sh.local_soft) {
      PF_LAUNCH(ctx, "orient_advect");
      launch_local_advect(st, scenes, ltiles, nltiles);
      launch_shape_match(st, scenes, btiles, nbtiles, p, c, 0, 0);
      nk += 2;
    }
    const int touch_valid = sh.iterations > 0 ? 1 : 0;
    bool friction_done = !apply;
    for (uint32_t f = 0; f < sh.nfluids; ++f) {
      const bool vort = (sh.fx[f] & PF_FX_VORTICITY) != 0;
      const bool visc = (sh.fx[f] & PF_FX_VISCOSITY) != 0;
      const bool lastf = f + 1 == sh.nfluids;
      if (vort || visc) {
        PF_LAUNCH(ctx, "velocity");
        launch_velocity(st, scenes, tiles, ntiles, (int)f, p, vort, visc, sh.any_boundary, sh.tu_cap[f]);
        ++nk;
      }
      // Friction follows every material's post-solve (loop.ts:241): it fuses
      // into the last fluid's vorticity pass unless a viscosity solver runs after it.
      const bool fric = lastf && sh.walls && sh.isv[f] == 0u && !sh.contact;
      if (vort || visc || fric) {
        PF_LAUNCH(ctx, fric ? "vorticity_friction" : "vorticity");
        launch_vorticity(st, scenes, tiles, ntiles, (int)f, p, c, vort, visc, fric, (int)k, touch_valid,
                         sh.tu_cap[f]);
        ++nk;
        if (fric) friction_done = true;
      }
      if (sh.isv[f]) {
        PF_LAUNCH(ctx, "viscosity_solve");
        launch_viscosity_solve(st, scenes, tiles, ntiles, (int)f, p, sh.isv[f]);
        nk += sh.isv[f] + 2u;
      }
    }
    if (!friction_done) {
      if (sh.contact && touch_valid) {
        PF_LAUNCH(ctx, "contact_friction");
        launch_contact_friction(st, scenes, ctiles, nctiles, p);
        ++nk;
      }
      PF_LAUNCH(ctx, "friction");
      if (touch_valid) launch_friction(st, scenes, tiles, ntiles, p, c, (int)k);
      ++nk;
    }
  }
  return nk;
}

// Parity bookkeeping of one step: returns the descriptor variant flips.
int step_variant_flips(const PlanShape& sh) {
  int state = 0, xs = 0;
  const bool grid = sh.grid;
  for (uint32_t k = 0; k < sh.substeps; ++k) {
    if (grid) state ^= 1;
    xs ^= 1;  // hash/predict
    if (grid && sh.contact) xs ^= 1;  // contact stabilisation
    for (uint32_t f = 0; f < sh.nfluids; ++f)
      if (sh.fx[f] & (PF_FX_SURFACE_TENSION | PF_FX_ADHESION)) xs ^= 1;
    for (uint32_t it = 0; it < sh.iterations; ++it)
      for (uint32_t f = 0; f < sh.nfluids; ++f) xs ^= 1;
  }
  return (state << 1) | xs;
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

static void alive(pf_particles* ps, const char* owner) {
  if (!ps) fail(PF_EINVAL, "%s: null handle", owner);
  if (ps->disposed) fail(PF_ESTATE, "%s: the ParticleSystem has been disposed", owner);
  if (ps->server) fail(PF_ESTATE, "%s: the scene is being served; remove it from the server first", owner);
}

static void range_ok(pf_particles* ps, uint32_t start, uint32_t count, const char* owner) {
  if ((uint64_t)start + count > ps->n)
    fail(PF_EINVAL, "%s: invalid particle range start=%u count=%u (capacity %u)", owner, start, count, ps->n);
}

// ---------------------------------------------------------------------------
// Quaternion helpers (double, host) for PrimitiveSet poses
// ---------------------------------------------------------------------------
struct Q {
  double x, y, z, w;
};
static Q qmul(Q a, Q b) {
  return Q{a.w * b.x + a.x * b.w + a.y * b.z - a.z * b.y, a.w * b.y - a.x * b.z + a.y * b.w + a.z * b.x,
           a.w * b.z + a.x * b.y - a.y * b.x + a.z * b.w, a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z};
}
static Q qinv(Q a) { return Q{-a.x, -a.y, -a.z, a.w}; }
static void qapply(Q q, const double v[3], double out[3]) {
  const double ix = q.w * v[0] + q.y * v[2] - q.z * v[1];
  const double iy = q.w * v[1] + q.z * v[0] - q.x * v[2];
  const double iz = q.w * v[2] + q.x * v[1] - q.y * v[0];
  const double iw = -q.x * v[0] - q.y * v[1] - q.z * v[2];
  out[0] = ix * q.w + iw * -q.x + iy * -q.z - iz * -q.y;
  out[1] = iy * q.w + iw * -q.y + iz * -q.x - ix * -q.z;
  out[2] = iz * q.w + iw * -q.z + ix * -q.y - iy * -q.x;
}

extern "C" {

// ---------------------------------------------------------------------------
// FluidSystem
// ---------------------------------------------------------------------------
int pf_fluid_create(pf_particles* ps, const pf_fluid_options* o, pf_fluid** out) {
  PF_API_BEGIN
  alive(ps, "FluidSystem");
  if (!out) fail(PF_EINVAL, "FluidSystem: null output");
  pf_fluid_options def{};
  def.compliance = -1;
  const pf_fluid_options& opt = o ? *o : def;
  uint32_t start = opt.range_start, count = opt.range_count;
  if (count == 0) { start = 0; count = ps->n; }
  range_ok(ps, start, count, "FluidSystem");
  const double rho0 = opt.rest_density > 0 ? opt.rest_density : 1000.0;
  if (!std::isfinite(rho0)) fail(PF_EINVAL, "FluidSystem: restDensity must be positive, got %g", rho0);
  const double spacing = opt.particle_spacing > 0 ? opt.particle_spacing : 2 * ps->radius;
  if (!std::isfinite(spacing)) fail(PF_EINVAL, "FluidSystem: particleSpacing must be positive, got %g", spacing);
  const double h = opt.smoothing_radius > 0 ? opt.smoothing_radius : 2 * spacing;
  if (!std::isfinite(h)) fail(PF_EINVAL, "FluidSystem: smoothingRadius must be positive, got %g", h);
  const double compliance = opt.compliance >= 0 ? opt.compliance : 1e-4;
  if (!std::isfinite(compliance)) fail(PF_EINVAL, "FluidSystem: compliance must be ≥ 0, got %g", compliance);
  const double vals[4] = {opt.viscosity, opt.vorticity, opt.surface_tension, opt.adhesion};
  const char* names[4] = {"viscosity", "vorticity", "surfaceTension", "adhesion"};
  for (int k = 0; k < 4; ++k)
    if ((opt.present >> k) & 1u)
      if (!std::isfinite(vals[k])) fail(PF_EINVAL, "FluidSystem: %s must be finite, got %g", names[k], vals[k]);
  auto* f = new pf_fluid_s();
  f->ps = ps;
  f->start = start;
  f->count = count;
  f->rest_density = rho0;
  f->spacing = spacing;
  f->h = h;
  f->compliance = compliance;
  f->mass = rho0 * spacing * spacing * spacing;  // FluidSystem.ts:109 (restDensity * spacing ** 3)
  f->present = opt.present & 15u;
  f->viscosity = opt.viscosity;
  f->vorticity = opt.vorticity;
  f->surface_tension = opt.surface_tension;
  f->adhesion = opt.adhesion;
  // The fluid owns its particles' mass (FluidSystem.ts:117-118) and its
  // density starts at the rest density (FluidSystem.ts:121-126).
  std::vector<float> im(count, (float)(1.0 / f->mass));
  for (uint32_t k = 0; k < count; ++k) ps->cpu_inv_mass[start + k] = im[k];
  particles_scatter(ps, start, count, nullptr, nullptr, &im, nullptr, 4u);
  std::vector<float> rho(count, (float)rho0);
  particles_scatter(ps, start, count, nullptr, nullptr, &rho, nullptr, 16u);
  *out = f;
  PF_API_END
}

static double* fluid_field(pf_fluid* f, uint32_t effect) {
  switch (effect) {
    case PF_FLUID_VISCOSITY: return &f->viscosity;
    case PF_FLUID_VORTICITY: return &f->vorticity;
    case PF_FLUID_SURFACE_TENSION: return &f->surface_tension;
    case PF_FLUID_ADHESION: return &f->adhesion;
  }
  return nullptr;
}

static const char* effect_name(uint32_t e) {
  switch (e) {
    case PF_FLUID_VISCOSITY: return "viscosity";
    case PF_FLUID_VORTICITY: return "vorticity";
    case PF_FLUID_SURFACE_TENSION: return "surfaceTension";
    case PF_FLUID_ADHESION: return "adhesion";
  }
  return "?";
}

static void fluid_write_params(pf_fluid* f);

int pf_fluid_set(pf_fluid* f, uint32_t effect, double value) {
  PF_API_BEGIN
  if (!f || f->disposed) fail(PF_ESTATE, "FluidSystem: disposed");
  double* fld = fluid_field(f, effect);
  if (!fld) fail(PF_EINVAL, "FluidSystem: unknown effect %u", effect);
  if (!(f->present & effect))
    fail(PF_EINVAL, "FluidSystem: pass `%s` in the options to enable it before changing it", effect_name(effect));
  if (!std::isfinite(value)) fail(PF_EINVAL, "FluidSystem: %s must be finite, got %g", effect_name(effect), value);
  *fld = value;
  if (f->loop) fluid_write_params(f);
  PF_API_END
}

int pf_fluid_get(pf_fluid* f, uint32_t effect, double* value) {
  PF_API_BEGIN
  if (!f || f->disposed) fail(PF_ESTATE, "FluidSystem: disposed");
  double* fld = fluid_field(f, effect);
  if (!fld) fail(PF_EINVAL, "FluidSystem: unknown effect %u", effect);
  if (!(f->present & effect))
    fail(PF_EINVAL, "FluidSystem: pass `%s` in the options to enable it before changing it", effect_name(effect));
  *value = *fld;
  PF_API_END
}

int pf_fluid_add_boundary(pf_fluid* f, uint32_t start, uint32_t count, int dynamic) {
  PF_API_BEGIN
  if (!f || f->disposed) fail(PF_ESTATE, "FluidSystem: disposed");
  if (f->loop) fail(PF_EINVAL, "FluidSystem.addBoundary: add boundaries before creating the SimLoop");
  range_ok(f->ps, start, count, "FluidSystem.addBoundary");
  auto overlaps = [](uint32_t a0, uint32_t ac, uint32_t b0, uint32_t bc) { return a0 < b0 + bc && b0 < a0 + ac; };
  if (overlaps(start, count, f->start, f->count)) fail(PF_EINVAL, "FluidSystem.addBoundary: a boundary cannot overlap the fluid");
  for (auto& b : f->boundaries)
    if (overlaps(start, count, b.start, b.count)) fail(PF_EINVAL, "FluidSystem.addBoundary: boundaries cannot overlap each other");
  if (f->boundaries.size() >= PF_MAX_BOUNDARIES) fail(PF_EINVAL, "FluidSystem.addBoundary: at most %d boundaries", PF_MAX_BOUNDARIES);
  f->boundaries.push_back({start, count, dynamic != 0});
  PF_API_END
}

int pf_fluid_readback_density(pf_fluid* f, float* density) {
  PF_API_BEGIN
  if (!f || f->disposed) fail(PF_ESTATE, "FluidSystem: disposed");
  if (f->ps->loop) PF_HIP(hipStreamSynchronize(f->ps->ctx->stream));
  std::vector<float> all(f->ps->n);
  particles_gather(f->ps, 4 /* density */, all.data(), sizeof(float));
  std::memcpy(density, all.data(), sizeof(float) * f->ps->n);
  PF_API_END
}

int pf_fluid_readback_overflow(pf_fluid* f, int* overflow) {
  PF_API_BEGIN
  if (!f || f->disposed) fail(PF_ESTATE, "FluidSystem: disposed");
  if (!f->loop) fail(PF_EINVAL, "FluidSystem: add the fluid to a SimLoop first");
  PF_HIP(hipStreamSynchronize(f->ps->ctx->stream));
  uint32_t fl[2 * PF_NUM_FLAGS];
  PF_HIP(hipMemcpy(fl, f->ps->host.flags, sizeof fl, hipMemcpyDeviceToHost));
  *overflow = fl[f->loop->last_flag_half * PF_NUM_FLAGS + PF_FLAG_NEIGHBORS] != 0;
  PF_API_END
}

int pf_fluid_info(pf_fluid* f, double* mass, double* h, double* spacing, double* rho0) {
  PF_API_BEGIN
  if (!f || f->disposed) fail(PF_ESTATE, "FluidSystem: disposed");
  if (mass) *mass = f->mass;
  if (h) *h = f->h;
  if (spacing) *spacing = f->spacing;
  if (rho0) *rho0 = f->rest_density;
  PF_API_END
}

// ---------------------------------------------------------------------------
// ViscositySolver (ViscositySolver.ts:38-131)
// ---------------------------------------------------------------------------
static void viscosity_ok(double v) {
  if (!(v >= 0) || !std::isfinite(v)) fail(PF_EINVAL, "ViscositySolver: viscosity must be ≥ 0, got %g", v);
}

int pf_viscosity_create(pf_fluid* f, double viscosity, uint32_t iterations, pf_viscosity** out) {
  PF_API_BEGIN
  if (!f || !out) fail(PF_EINVAL, "ViscositySolver: null argument");
  if (f->disposed) fail(PF_ESTATE, "ViscositySolver: the FluidSystem has been disposed");
  if (iterations < 1 || iterations > 64)
    fail(PF_EINVAL, "ViscositySolver: iterations must be an integer from 1 to 64, got %u", iterations);
  viscosity_ok(viscosity);
  auto* v = new pf_viscosity_s();
  v->fluid = f;
  v->viscosity = viscosity;
  v->iterations = iterations;
  *out = v;
  PF_API_END
}

static void fluid_write_params(pf_fluid* f);

int pf_viscosity_set(pf_viscosity* v, double viscosity) {
  PF_API_BEGIN
  if (!v || v->disposed) fail(PF_ESTATE, "ViscositySolver: disposed");
  if (v->fluid && v->fluid->ps && v->fluid->ps->server)
    fail(PF_ESTATE, "ViscositySolver: the scene is being served; remove it from the server first");
  viscosity_ok(viscosity);
  v->viscosity = viscosity;
  if (v->loop && !v->fluid->disposed) fluid_write_params(v->fluid);
  PF_API_END
}

int pf_viscosity_destroy(pf_viscosity* v) {
  PF_API_BEGIN
  if (!v || v->disposed) return PF_OK;
  v->disposed = true;
  if (v->loop) (void)hipStreamSynchronize(v->loop->ctx->stream);
  PF_API_END
}

static void loop_invalidate_graphs(pf_loop* L);

int pf_fluid_set_tile_union_limit(pf_fluid* f, uint32_t max_particles) {
  PF_API_BEGIN
  if (!f || f->disposed) fail(PF_ESTATE, "FluidSystem: disposed");
  if (f->ps && f->ps->server) fail(PF_ESTATE, "FluidSystem: the scene is being served; remove it from the server first");
  if (max_particles > PF_TU_MAX) fail(PF_EINVAL, "FluidSystem: the tile union limit is at most %u, got %u", PF_TU_MAX, max_particles);
  f->tu_limit = max_particles;
  if (f->loop) {
    fluid_write_params(f);
    // captured launches carry the LDS capacity the limit implies
    loop_invalidate_graphs(f->loop);
  }
  PF_API_END
}

int pf_fluid_tile_union_stats(pf_fluid* f, uint32_t* tiles, uint32_t* global_tiles, uint32_t* largest,
                              double* mean) {
  PF_API_BEGIN
  if (!f || f->disposed) fail(PF_ESTATE, "FluidSystem: disposed");
  if (!f->loop || !f->tu_size) fail(PF_EINVAL, "FluidSystem: add the fluid to a SimLoop and step it first");
  PF_HIP(hipStreamSynchronize(f->ps->ctx->stream));
  const size_t ntiles = ((size_t)f->ps->n + PF_BLOCK - 1) / PF_BLOCK;
  std::vector<uint32_t> u(ntiles);
  PF_HIP(hipMemcpy(u.data(), f->tu_size, sizeof(uint32_t) * ntiles, hipMemcpyDeviceToHost));
  uint32_t g = 0, big = 0;
  double sum = 0;
  for (uint32_t v : u) {
    if (v == PF_TU_GLOBAL) {
      ++g;
    } else {
      big = std::max(big, v);
      sum += v;
    }
  }
  if (tiles) *tiles = (uint32_t)ntiles;
  if (global_tiles) *global_tiles = g;
  if (largest) *largest = big;
  if (mean) *mean = ntiles > g ? sum / (double)(ntiles - g) : 0.0;
  PF_API_END
}

int pf_fluid_destroy(pf_fluid* f) {
  PF_API_BEGIN
  if (!f || f->disposed) return PF_OK;
  f->disposed = true;
  dfree(f->nl); dfree(f->nl_count); dfree(f->normal); dfree(f->omega); dfree(f->dv);
  dfree(f->reaction);
  dfree(f->nl16); dfree(f->tu_size); dfree(f->tu_slots);
  f->nl16 = f->tu_size = f->tu_slots = nullptr;
  f->nl = nullptr; f->nl_count = nullptr; f->normal = f->omega = f->dv = nullptr;
  f->reaction = nullptr;
  PF_API_END
}

// ---------------------------------------------------------------------------
// PrimitiveSet
// ---------------------------------------------------------------------------
int pf_primitives_create(pf_particles* ps, uint32_t capacity, pf_primitives** out) {
  PF_API_BEGIN
  alive(ps, "PrimitiveSet");
  auto* s = new pf_primitives_s();
  s->ps = ps;
  s->reserved = capacity;
  *out = s;
  PF_API_END
}

static void friction_defaults(double& mus, double& muk, const char* owner) {
  if (mus < 0) mus = 0.5;
  if (muk < 0) muk = 0.4;
  if (!(mus >= 0) || !(muk >= 0) || !std::isfinite(mus) || !std::isfinite(muk))
    fail(PF_EINVAL, "%s: friction coefficients must be non-negative (got muS=%g, muK=%g)", owner, mus, muk);
}

static uint32_t prim_add(pf_primitives* s, uint32_t kind, uint32_t flags, const double d0[4], const double d1[4],
                         double mus, double muk, const double* rot) {
  if (!s || s->disposed) fail(PF_ESTATE, "PrimitiveSet has been disposed");
  const uint32_t cap = s->loop ? s->capacity : (s->reserved ? s->reserved : 0);
  if (cap && s->prims.size() >= cap)
    fail(PF_EINVAL, "PrimitiveSet: capacity %u is full. Pass a larger `capacity` to reserve slots for primitives added after the simulation starts.", cap);
  if (s->prims.size() >= 32) fail(PF_EINVAL, "PrimitiveSet: at most 32 primitives per set");
  friction_defaults(mus, muk, "PrimitiveSet");
  pf_primitives_s::Prim p{};
  p.kind = kind;
  p.flags = flags;
  for (int a = 0; a < 4; ++a) { p.d0[a] = d0[a]; p.d1[a] = d1[a]; }
  p.d1[3] = mus;
  p.vel[0] = p.vel[1] = p.vel[2] = 0;
  p.vel[3] = muk;
  p.spin[0] = p.spin[1] = p.spin[2] = 0;
  p.rot[0] = p.rot[1] = p.rot[2] = 0;
  p.rot[3] = 1;
  if (rot) for (int a = 0; a < 4; ++a) p.rot[a] = rot[a];
  s->prims.push_back(p);
  s->dirty = true;
  return (uint32_t)s->prims.size() - 1;
}

int pf_primitives_add_plane(pf_primitives* s, const double n[3], const double pt[3], double mus, double muk,
                            uint32_t* slot) {
  PF_API_BEGIN
  const double l2 = n[0] * n[0] + n[1] * n[1] + n[2] * n[2];
  if (!(l2 > 0)) fail(PF_EINVAL, "PrimitiveSet.addPlane: normal must be non-zero");
  // Vector3.normalize(): multiply by 1 / length.
  const double inv = 1.0 / std::sqrt(l2);
  const double d0[4] = {n[0] * inv, n[1] * inv, n[2] * inv, 0};
  const double d1[4] = {pt[0], pt[1], pt[2], 0};
  const uint32_t k = prim_add(s, PF_KIND_PLANE, 0, d0, d1, mus, muk, nullptr);
  if (slot) *slot = k;
  PF_API_END
}

int pf_primitives_add_sphere(pf_primitives* s, const double c[3], double r, int invert, double mus, double muk,
                             uint32_t* slot) {
  PF_API_BEGIN
  if (!(r > 0) || !std::isfinite(r)) fail(PF_EINVAL, "PrimitiveSet.addSphere: radius must be positive, got %g", r);
  const double d0[4] = {c[0], c[1], c[2], r};
  const double d1[4] = {0, 0, 0, 0};
  const uint32_t k = prim_add(s, PF_KIND_SPHERE, invert ? PF_PRIM_INVERT : 0u, d0, d1, mus, muk, nullptr);
  if (slot) *slot = k;
  PF_API_END
}

int pf_primitives_add_box(pf_primitives* s, const double c[3], const double he[3], const double rot[4], int invert,
                          double mus, double muk, uint32_t* slot) {
  PF_API_BEGIN
  if (!(he[0] > 0 && he[1] > 0 && he[2] > 0)) fail(PF_EINVAL, "PrimitiveSet.addBox: halfExtents must be positive");
  double q[4] = {0, 0, 0, 1};
  if (rot) {
    const double l = std::sqrt(rot[0] * rot[0] + rot[1] * rot[1] + rot[2] * rot[2] + rot[3] * rot[3]);
    if (!(l > 0) || !std::isfinite(l)) fail(PF_EINVAL, "PrimitiveSet.addBox: rotation must be a finite, non-zero quaternion");
    for (int a = 0; a < 4; ++a) q[a] = rot[a] / l;
  }
  const double d0[4] = {c[0], c[1], c[2], 0};
  const double d1[4] = {he[0], he[1], he[2], 0};
  const uint32_t k = prim_add(s, PF_KIND_BOX, invert ? PF_PRIM_INVERT : 0u, d0, d1, mus, muk, q);
  if (slot) *slot = k;
  PF_API_END
}

int pf_primitives_add_capsule(pf_primitives* s, const double a[3], const double b[3], double r, double mus,
                              double muk, uint32_t* slot) {
  PF_API_BEGIN
  if (!(r > 0) || !std::isfinite(r)) fail(PF_EINVAL, "PrimitiveSet.addCapsule: radius must be positive, got %g", r);
  const double d0[4] = {a[0], a[1], a[2], r};
  const double d1[4] = {b[0], b[1], b[2], 0};
  const uint32_t k = prim_add(s, PF_KIND_CAPSULE, 0u, d0, d1, mus, muk, nullptr);
  if (slot) *slot = k;
  PF_API_END
}

static pf_primitives_s::Prim& prim_at(pf_primitives* s, uint32_t slot) {
  if (!s || s->disposed) fail(PF_ESTATE, "PrimitiveSet has been disposed");
  if (slot >= s->prims.size()) fail(PF_EINVAL, "PrimitiveSet: no primitive in slot %u", slot);
  return s->prims[slot];
}

int pf_primitives_set_sphere(pf_primitives* s, uint32_t slot, const double c[3], double r, const double v[3]) {
  PF_API_BEGIN
  auto& p = prim_at(s, slot);
  if (p.kind != PF_KIND_SPHERE) fail(PF_EINVAL, "PrimitiveSet.setSphere: slot %u is not a sphere", slot);
  if (!(r > 0) || !std::isfinite(r)) fail(PF_EINVAL, "PrimitiveSet.setSphere: radius must be positive, got %g", r);
  p.d0[0] = c[0]; p.d0[1] = c[1]; p.d0[2] = c[2]; p.d0[3] = r;
  p.flags |= PF_PRIM_SWEEP;
  p.vel[0] = v ? v[0] : 0; p.vel[1] = v ? v[1] : 0; p.vel[2] = v ? v[2] : 0;
  s->dirty = true;
  PF_API_END
}

// place() (PrimitiveSet.ts:469-486)
static void prim_place(pf_primitives_s::Prim& p, const double pos[3], Q rot) {
  const Q r = qmul(rot, Q{p.offset[0], p.offset[1], p.offset[2], p.offset[3]});
  p.rot[0] = r.x; p.rot[1] = r.y; p.rot[2] = r.z; p.rot[3] = r.w;
  double d[3];
  qapply(r, p.direction, d);
  if (p.kind == PF_KIND_PLANE) {
    p.d0[0] = d[0]; p.d0[1] = d[1]; p.d0[2] = d[2];
    p.d1[0] = pos[0]; p.d1[1] = pos[1]; p.d1[2] = pos[2];
  } else if (p.kind == PF_KIND_CAPSULE) {
    for (int a = 0; a < 3; ++a) { p.d0[a] = pos[a] - d[a]; p.d1[a] = pos[a] + d[a]; }
  } else {
    p.d0[0] = pos[0]; p.d0[1] = pos[1]; p.d0[2] = pos[2];
  }
}

int pf_primitives_set_pose(pf_primitives* s, uint32_t slot, const double pos[3], const double rot[4], double dt) {
  PF_API_BEGIN
  auto& p = prim_at(s, slot);
  const Q q = rot ? Q{rot[0], rot[1], rot[2], rot[3]} : Q{0, 0, 0, 1};
  if (!p.has_pose) {
    // attach() (PrimitiveSet.ts:211-234)
    if (p.kind == PF_KIND_PLANE) { p.direction[0] = p.d0[0]; p.direction[1] = p.d0[1]; p.direction[2] = p.d0[2]; }
    else if (p.kind == PF_KIND_CAPSULE) for (int a = 0; a < 3; ++a) p.direction[a] = (p.d1[a] - p.d0[a]) * 0.5;
    else p.direction[0] = p.direction[1] = p.direction[2] = 0;
    const Q off = p.kind == PF_KIND_BOX ? Q{0, 0, 0, 1} : qinv(q);
    p.offset[0] = off.x; p.offset[1] = off.y; p.offset[2] = off.z; p.offset[3] = off.w;
    p.flags |= PF_PRIM_SWEEP;
    p.has_pose = true;
  } else if (dt > 0) {
    // update(dt) (PrimitiveSet.ts:259-285)
    for (int a = 0; a < 3; ++a) p.vel[a] = (pos[a] - p.prev_pos[a]) / dt;
    Q turn = qmul(q, qinv(Q{p.prev_rot[0], p.prev_rot[1], p.prev_rot[2], p.prev_rot[3]}));
    if (turn.w < 0) turn = Q{-turn.x, -turn.y, -turn.z, -turn.w};
    const double sn = std::hypot(turn.x, std::hypot(turn.y, turn.z));
    const double rate = sn > 1e-12 ? (2 * std::atan2(sn, turn.w)) / (sn * dt) : 0;
    p.spin[0] = turn.x * rate; p.spin[1] = turn.y * rate; p.spin[2] = turn.z * rate;
  }
  prim_place(p, pos, q);
  for (int a = 0; a < 3; ++a) p.prev_pos[a] = pos[a];
  p.prev_rot[0] = q.x; p.prev_rot[1] = q.y; p.prev_rot[2] = q.z; p.prev_rot[3] = q.w;
  s->dirty = true;
  PF_API_END
}

int pf_primitives_count(pf_primitives* s, uint32_t* count) {
  PF_API_BEGIN
  if (!s || s->disposed) fail(PF_ESTATE, "PrimitiveSet has been disposed");
  *count = (uint32_t)s->prims.size();
  PF_API_END
}

int pf_primitives_destroy(pf_primitives* s) {
  PF_API_BEGIN
  if (!s || s->disposed) return PF_OK;
  s->disposed = true;
  dfree(s->packed); dfree(s->d0); dfree(s->d1); dfree(s->vel); dfree(s->rot); dfree(s->spin);
  s->packed = nullptr; s->d0 = s->d1 = s->vel = s->rot = s->spin = nullptr;
  PF_API_END
}

// ---------------------------------------------------------------------------
// SDFCollider (SDFCollider.ts:68-311)
// ---------------------------------------------------------------------------
static void sdf_friction_ok(double mus, double muk) {
  if (!(mus >= 0) || !(muk >= 0))
    fail(PF_EINVAL, "SDFCollider: friction coefficients must be non-negative (got muS=%g, muK=%g)", mus, muk);
}

static void sdf_alive(pf_sdf_collider* c) {
  if (!c) fail(PF_EINVAL, "SDFCollider: null handle");
  if (c->disposed) fail(PF_ESTATE, "SDFCollider: disposed");
}

// Placement changes reach the scene descriptor of the loop that holds it.
static void sdf_changed(pf_sdf_collider* c) {
  if (c->loop && c->slot >= 0) {
    c->write(c->ps->host.sdf[c->slot]);
    c->ps->dirty = true;
  }
}

static void sdf_set_quat(pf_sdf_collider* c, const double q[4], const char* owner) {
  const double l = std::sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
  if (!(l > 0) || !std::isfinite(l)) fail(PF_EINVAL, "%s: rotation must be a finite, non-zero quaternion", owner);
  const double inv = 1.0 / l;  // Quaternion.normalize()
  for (int a = 0; a < 4; ++a) c->quat[a] = q[a] * inv;
}

int pf_sdf_collider_create(pf_particles* ps, const float* data, const uint32_t res[3], const double origin[3],
                           const double voxel[3], const pf_sdf_options* opt, pf_sdf_collider** out) {
  PF_API_BEGIN
  alive(ps, "SDFCollider");
  ArenaScope arena_scope(ps->arena.get());
  if (!(data && res && origin && voxel && out)) fail(PF_EINVAL, "SDFCollider: null argument");
  for (int a = 0; a < 3; ++a)
    if (res[a] < 4) fail(PF_EINVAL, "SDFCollider: resolution must be integers ≥ 4, got [%u, %u, %u]", res[0], res[1], res[2]);
  for (int a = 0; a < 3; ++a)
    if (!(voxel[a] > 0) || !std::isfinite(voxel[a])) fail(PF_EINVAL, "SDFCollider: voxelSize must be positive");
  auto c = std::make_unique<pf_sdf_collider_s>();
  c->ps = ps;
  for (int a = 0; a < 3; ++a) {
    c->res[a] = res[a];
    c->origin[a] = origin[a];
    c->voxel[a] = voxel[a];
  }
  if (opt) {
    const double mus = opt->mu_s < 0 ? 0.5 : opt->mu_s, muk = opt->mu_k < 0 ? 0.4 : opt->mu_k;
    sdf_friction_ok(mus, muk);
    c->mus = mus;
    c->muk = muk;
    if (!(opt->thickness >= 0) || !std::isfinite(opt->thickness))
      fail(PF_EINVAL, "SDFCollider: thickness must be non-negative, got %g", opt->thickness);
    c->thickness = opt->thickness;
    for (int a = 0; a < 3; ++a) c->pos[a] = opt->position[a];
    if (opt->rotation) sdf_set_quat(c.get(), opt->rotation, "SDFCollider.setRotation");
    if (opt->scale > 0) {
      if (!std::isfinite(opt->scale)) fail(PF_EINVAL, "SDFCollider.setScale: scale must be positive, got %g", opt->scale);
      c->scale = opt->scale;
    }
  }
  const size_t nv = (size_t)res[0] * res[1] * res[2];
  std::vector<uint16_t> half(nv);
  for (size_t k = 0; k < nv; ++k) half[k] = half_from_float(data[k]);
  c->grid = dalloc<uint16_t>(nv);
  PF_HIP(hipMemcpy(c->grid, half.data(), nv * sizeof(uint16_t), hipMemcpyHostToDevice));
  *out = c.release();
  PF_API_END
}

int pf_sdf_collider_set_position(pf_sdf_collider* c, const double p[3]) {
  PF_API_BEGIN
  sdf_alive(c);
  for (int a = 0; a < 3; ++a) c->pos[a] = p[a];
  sdf_changed(c);
  PF_API_END
}

int pf_sdf_collider_set_rotation(pf_sdf_collider* c, const double q[4]) {
  PF_API_BEGIN
  sdf_alive(c);
  sdf_set_quat(c, q, "SDFCollider.setRotation");
  sdf_changed(c);
  PF_API_END
}

int pf_sdf_collider_set_scale(pf_sdf_collider* c, double scale) {
  PF_API_BEGIN
  sdf_alive(c);
  if (!(scale > 0) || !std::isfinite(scale)) fail(PF_EINVAL, "SDFCollider.setScale: scale must be positive, got %g", scale);
  c->scale = scale;
  sdf_changed(c);
  PF_API_END
}

int pf_sdf_collider_set_friction(pf_sdf_collider* c, double mus, double muk) {
  PF_API_BEGIN
  sdf_alive(c);
  sdf_friction_ok(mus, muk);
  c->mus = mus;
  c->muk = muk;
  sdf_changed(c);
  PF_API_END
}

int pf_sdf_collider_set_thickness(pf_sdf_collider* c, double t) {
  PF_API_BEGIN
  sdf_alive(c);
  if (!(t >= 0) || !std::isfinite(t)) fail(PF_EINVAL, "SDFCollider: thickness must be non-negative, got %g", t);
  c->thickness = t;
  sdf_changed(c);
  PF_API_END
}

int pf_sdf_collider_destroy(pf_sdf_collider* c) {
  PF_API_BEGIN
  if (!c || c->disposed) return PF_OK;
  if (c->loop) fail(PF_ESTATE, "SDFCollider: dispose the SimLoop that uses it first");
  c->disposed = true;
  dfree(c->grid);
  c->grid = nullptr;
  PF_API_END
}

}  // extern "C"

// The uniforms SDFCollider keeps (SDFCollider.ts:115-139): rotation from the
// quaternion as Matrix4.makeRotationFromQuaternion, its transpose, 1/scale,
// 1/voxel and 1/resolution, all formed in double and stored as f32.
void pf_sdf_collider_s::write(SdfDev& D) const {
  D.grid = grid;
  D.nx = (int)res[0];
  D.ny = (int)res[1];
  D.nz = (int)res[2];
  for (int a = 0; a < 3; ++a) {
    D.res[a] = (float)res[a];
    D.inv_res[a] = (float)(1.0 / (double)res[a]);
    D.origin[a] = (float)origin[a];
    D.voxel[a] = (float)voxel[a];
    D.inv_voxel[a] = (float)(1.0 / voxel[a]);
    D.pos[a] = (float)pos[a];
  }
  const double x = quat[0], y = quat[1], z = quat[2], w = quat[3];
  const double x2 = x + x, y2 = y + y, z2 = z + z;
  const double xx = x * x2, xy = x * y2, xz = x * z2, yy = y * y2, yz = y * z2, zz = z * z2;
  const double wx = w * x2, wy = w * y2, wz = w * z2;
  const double R[9] = {1 - (yy + zz), xy - wz, xz + wy, xy + wz, 1 - (xx + zz), yz - wx, xz - wy, yz + wx, 1 - (xx + yy)};
  for (int r = 0; r < 3; ++r)
    for (int cc = 0; cc < 3; ++cc) {
      D.rot[3 * r + cc] = (float)R[3 * r + cc];
      D.inv_rot[3 * r + cc] = (float)R[3 * cc + r];
    }
  D.scale = (float)scale;
  D.inv_scale = (float)(1.0 / scale);
  D.thickness = (float)thickness;
  D.mus = (float)mus;
  D.muk = (float)muk;
}

void pf_primitives_s::upload() {
  if (!dirty || !packed) return;
  const uint32_t n = (uint32_t)prims.size();
  std::vector<uint32_t> pk(capacity, 0);
  std::vector<float4> a(capacity), b(capacity), v(capacity), r(capacity), sp(capacity);
  for (uint32_t i = 0; i < n; ++i) {
    const Prim& p = prims[i];
    pk[i] = (p.kind << 16) | p.flags;
    a[i] = make_float4((float)p.d0[0], (float)p.d0[1], (float)p.d0[2], (float)p.d0[3]);
    b[i] = make_float4((float)p.d1[0], (float)p.d1[1], (float)p.d1[2], (float)p.d1[3]);
    v[i] = make_float4((float)p.vel[0], (float)p.vel[1], (float)p.vel[2], (float)p.vel[3]);
    r[i] = make_float4((float)p.rot[0], (float)p.rot[1], (float)p.rot[2], (float)p.rot[3]);
    sp[i] = make_float4((float)p.spin[0], (float)p.spin[1], (float)p.spin[2], 0.f);
  }
  PF_HIP(hipStreamSynchronize(ps->ctx->stream));
  PF_HIP(hipMemcpy(packed, pk.data(), capacity * 4, hipMemcpyHostToDevice));
  PF_HIP(hipMemcpy(d0, a.data(), capacity * 16, hipMemcpyHostToDevice));
  PF_HIP(hipMemcpy(d1, b.data(), capacity * 16, hipMemcpyHostToDevice));
  PF_HIP(hipMemcpy(vel, v.data(), capacity * 16, hipMemcpyHostToDevice));
  PF_HIP(hipMemcpy(rot, r.data(), capacity * 16, hipMemcpyHostToDevice));
  PF_HIP(hipMemcpy(spin, sp.data(), capacity * 16, hipMemcpyHostToDevice));
  if (ps->host.nprim != n) {
    ps->host.nprim = n;
    ps->dirty = true;
  }
  dirty = false;
}

extern "C" {

// ---------------------------------------------------------------------------
// SoftbodySystem (global shape matching)
// ---------------------------------------------------------------------------
// Smallest eigenvalue of a symmetric 3x3 by Jacobi (SoftbodySystem.ts:424-486).
static double smallest_eigenvalue(double a00, double a01, double a02, double a11, double a12, double a22) {
  for (int sweep = 0; sweep < 30; ++sweep) {
    const double off = std::fabs(a01) + std::fabs(a02) + std::fabs(a12);
    if (off < 1e-14) break;
    const int pivot = (std::fabs(a01) >= std::fabs(a02) && std::fabs(a01) >= std::fabs(a12)) ? 0
                      : (std::fabs(a02) >= std::fabs(a12) ? 1 : 2);
    const double apq = pivot == 0 ? a01 : pivot == 1 ? a02 : a12;
    const double app = pivot == 2 ? a11 : a00;
    const double aqq = pivot == 0 ? a11 : a22;
    if (std::fabs(apq) < 1e-20) break;
    const double theta = (aqq - app) / (2 * apq);
    const double sg = theta > 0 ? 1.0 : (theta < 0 ? -1.0 : 1.0);
    const double t = sg / (std::fabs(theta) + std::sqrt(1 + theta * theta));
    const double c = 1 / std::sqrt(1 + t * t);
    const double s = t * c;
    if (pivot == 0) {
      const double n02 = c * a02 - s * a12, n12 = s * a02 + c * a12;
      a00 = app - t * apq; a11 = aqq + t * apq; a01 = 0; a02 = n02; a12 = n12;
    } else if (pivot == 1) {
      const double n01 = c * a01 - s * a12, n12 = s * a01 + c * a12;
      a00 = app - t * apq; a22 = aqq + t * apq; a02 = 0; a01 = n01; a12 = n12;
    } else {
      const double n01 = c * a01 - s * a02, n02 = s * a01 + c * a02;
      a11 = app - t * apq; a22 = aqq + t * apq; a12 = 0; a01 = n01; a02 = n02;
    }
  }
  return std::min(a00, std::min(a11, a22));
}

int pf_softbody_create(pf_particles* ps, const pf_softbody_def* defs, uint32_t nb, uint32_t mode,
                       pf_softbody** out) {
  PF_API_BEGIN
  alive(ps, "SoftbodySystem");
  ArenaScope arena_scope(ps->arena.get());
  if (nb == 0 || !defs) fail(PF_EINVAL, "SoftbodySystem: at least one body is required");
  if (mode != PF_SHAPE_MATCHING_GLOBAL && mode != PF_SHAPE_MATCHING_LOCAL)
    fail(PF_EINVAL, "SoftbodySystem: shapeMatching must be 'global' or 'local'");
  auto sb = std::make_unique<pf_softbody_s>();
  sb->ps = ps;
  sb->mode = mode;
  const uint32_t n = ps->n;
  sb->rest.assign(4 * (size_t)n, 0.0f);
  sb->weights.assign(n, 0.0f);
  std::vector<uint8_t> used(n, 0);
  for (uint32_t b = 0; b < nb; ++b) {
    const pf_softbody_def& d = defs[b];
    if (d.count == 0) fail(PF_EINVAL, "SoftbodySystem: body %u has no particles", b);
    char owner[64];
    std::snprintf(owner, sizeof owner, "SoftbodySystem body %u", b);
    range_ok(ps, d.start, d.count, owner);
    for (uint32_t i = d.start; i < d.start + d.count; ++i) {
      if (used[i]) fail(PF_EINVAL, "SoftbodySystem: particle %u belongs to two bodies", i);
      used[i] = 1;
    }
    if (d.rest_positions && (uint64_t)d.rest_count != 3ull * d.count)
      fail(PF_EINVAL, "SoftbodySystem: body %u needs %llu rest coordinates", b, 3ull * d.count);
    std::vector<float> restp(3 * (size_t)d.count);
    for (uint32_t k = 0; k < 3 * d.count; ++k)
      restp[k] = d.rest_positions ? d.rest_positions[k] : ps->cpu_positions[3 * (size_t)d.start + k];
    const uint32_t surface = d.surface_count == 0xffffffffu ? d.count : d.surface_count;
    if (surface > d.count) fail(PF_EINVAL, "SoftbodySystem: body %u surfaceCount %u is out of range", b, surface);
    if (d.edges) {
      // assertEdges (SoftbodySystem.ts:453-463)
      for (uint32_t e = 0; e < d.edge_count; ++e) {
        const uint32_t i = d.edges[2 * e], j = d.edges[2 * e + 1];
        if (!(i < d.count) || !(j < d.count) || i == j)
          fail(PF_EINVAL, "SoftbodySystem: body %u has an invalid edge (%u, %u)", b, i, j);
      }
    } else if (mode == PF_SHAPE_MATCHING_LOCAL) {
      fail(PF_EINVAL, "SoftbodySystem: body %u needs edges for local shape matching", b);
    }
    if (!(d.compliance >= 0) || !std::isfinite(d.compliance))
      fail(PF_EINVAL, "SoftbodySystem: compliance must be ≥ 0, got %g", d.compliance);
    // massWeights (SoftbodySystem.ts:366-372)
    std::vector<double> masses(d.count);
    std::vector<double> dyn;
    for (uint32_t i = 0; i < d.count; ++i) {
      const float w = ps->cpu_inv_mass[d.start + i];
      masses[i] = w > 0 ? 1.0 / (double)w : 0.0;
      if (masses[i] > 0) dyn.push_back(masses[i]);
    }
    bool equal = true;
    for (double m : dyn) equal = equal && (m == dyn[0]);
    std::vector<float> wts(d.count, 1.0f);
    if (!equal) {
      double sum = 0;
      for (double m : dyn) sum += m;
      const double mean = sum / (double)dyn.size();
      for (uint32_t i = 0; i < d.count; ++i) wts[i] = (float)(masses[i] > 0 ? masses[i] / mean : 1.0);
    }
    // centroid (SoftbodySystem.ts:374-386)
    double cx = 0, cy = 0, cz = 0, tot = 0;
    for (uint32_t i = 0; i < d.count; ++i) {
      cx += (double)restp[3 * i] * wts[i];
      cy += (double)restp[3 * i + 1] * wts[i];
      cz += (double)restp[3 * i + 2] * wts[i];
      tot += wts[i];
    }
    cx /= tot; cy /= tot; cz /= tot;
    // assertVolumetric (SoftbodySystem.ts:392-421)
    double xx = 0, xy = 0, xz = 0, yy = 0, yz = 0, zz = 0;
    for (uint32_t i = 0; i < d.count; ++i) {
      const double x = restp[3 * i] - cx, y = restp[3 * i + 1] - cy, z = restp[3 * i + 2] - cz;
      xx += x * x; xy += x * y; xz += x * z; yy += y * y; yz += y * z; zz += z * z;
    }
    const double smallest = std::sqrt(std::max(0.0, smallest_eigenvalue(xx, xy, xz, yy, yz, zz)));
    if (smallest < ps->radius * 1e-3)
      fail(PF_EINVAL, "SoftbodySystem: body %u is flat; shape matching needs a 3D particle cloud", b);
    for (uint32_t i = 0; i < d.count; ++i) {
      const size_t g = d.start + i;
      sb->rest[4 * g + 0] = (float)(restp[3 * i] - cx);
      sb->rest[4 * g + 1] = (float)(restp[3 * i + 1] - cy);
      sb->rest[4 * g + 2] = (float)(restp[3 * i + 2] - cz);
      sb->rest[4 * g + 3] = (float)d.compliance;
      sb->weights[g] = wts[i];
    }
    sb->bodies.push_back({d.start, d.count, surface, d.compliance});
  }
  const uint32_t B = nb;
  std::vector<uint32_t> bs(B), bc(B);
  std::vector<float> cmp(B);
  for (uint32_t b = 0; b < B; ++b) { bs[b] = sb->bodies[b].start; bc[b] = sb->bodies[b].count; cmp[b] = (float)sb->bodies[b].compliance; }
  sb->d_start = dalloc<uint32_t>(B);
  sb->d_count = dalloc<uint32_t>(B);
  sb->d_compliance = dalloc<float>(B);
  sb->d_rest = dalloc<float4>(n);
  sb->d_weights = dalloc<float>(n);
  sb->d_lambda = dalloc<float4>(n);
  sb->d_centers = dalloc<float4>(B);
  sb->d_rotations = dalloc<float4>(3 * (size_t)B);
  PF_HIP(hipMemcpy(sb->d_start, bs.data(), B * 4, hipMemcpyHostToDevice));
  PF_HIP(hipMemcpy(sb->d_count, bc.data(), B * 4, hipMemcpyHostToDevice));
  PF_HIP(hipMemcpy(sb->d_compliance, cmp.data(), B * 4, hipMemcpyHostToDevice));
  PF_HIP(hipMemcpy(sb->d_rest, sb->rest.data(), (size_t)n * 16, hipMemcpyHostToDevice));
  PF_HIP(hipMemcpy(sb->d_weights, sb->weights.data(), (size_t)n * 4, hipMemcpyHostToDevice));
  // Identity rotation until the first step (SoftbodySystem.ts:147-151).
  std::vector<float4> rot(3 * (size_t)B);
  for (uint32_t b = 0; b < B; ++b) {
    rot[3 * b] = make_float4(1, 0, 0, 0); rot[3 * b + 1] = make_float4(0, 1, 0, 0); rot[3 * b + 2] = make_float4(0, 0, 1, 0);
  }
  PF_HIP(hipMemcpy(sb->d_rotations, rot.data(), rot.size() * 16, hipMemcpyHostToDevice));
  if (mode == PF_SHAPE_MATCHING_LOCAL) {
    // buildNeighborGraph (SoftbodySystem.ts:470-495): row i = i, then its edge
    // neighbours in edge order, over global ids.
    std::vector<uint32_t> deg(n, 0);
    uint32_t lo = 0xffffffffu, hi = 0;
    for (uint32_t b = 0; b < nb; ++b) {
      const pf_softbody_def& d = defs[b];
      for (uint32_t i = 0; i < d.count; ++i) deg[d.start + i] += 1;
      for (uint32_t e = 0; e < d.edge_count; ++e) {
        deg[d.start + d.edges[2 * e]] += 1;
        deg[d.start + d.edges[2 * e + 1]] += 1;
      }
      lo = std::min(lo, d.start);
      hi = std::max(hi, d.start + d.count);
    }
    std::vector<uint32_t> off(n + 1, 0);
    for (uint32_t i = 0; i < n; ++i) off[i + 1] = off[i] + deg[i];
    std::vector<uint32_t> idx(std::max<uint32_t>(1, off[n]), 0);
    std::vector<uint32_t> cur(off.begin(), off.end() - 1);
    for (uint32_t b = 0; b < nb; ++b) {
      const pf_softbody_def& d = defs[b];
      for (uint32_t i = d.start; i < d.start + d.count; ++i) idx[cur[i]++] = i;
      for (uint32_t e = 0; e < d.edge_count; ++e) {
        const uint32_t i = d.start + d.edges[2 * e], j = d.start + d.edges[2 * e + 1];
        idx[cur[i]++] = j;
        idx[cur[j]++] = i;
      }
    }
    // Rest centre of each group, accumulated into f32 storage one member at
    // a time like the reference's Float32Array (SoftbodySystem.ts:281-292).
    std::vector<float4> rc(n, make_float4(0.f, 0.f, 0.f, 0.f));
    for (uint32_t b = 0; b < nb; ++b) {
      const pf_softbody_def& d = defs[b];
      for (uint32_t i = d.start; i < d.start + d.count; ++i) {
        const uint32_t from = off[i], to = off[i + 1];
        float* acc = &rc[i].x;
        for (uint32_t k = from; k < to; ++k) {
          const uint32_t j = idx[k];
          for (int a = 0; a < 3; ++a)
            acc[a] = (float)((double)acc[a] + (double)sb->rest[4 * (size_t)j + a] / (double)(to - from));
        }
      }
    }
    sb->span_lo = lo;
    sb->span_hi = hi;
    sb->d_csr_off = dalloc<uint32_t>(n + 1);
    sb->d_csr_idx = dalloc<uint32_t>(idx.size());
    sb->d_rest_center = dalloc<float4>(n);
    sb->d_group_center = dalloc<float4>(n);
    sb->d_group_rot = dalloc<float4>(3 * (size_t)n);
    sb->d_pair_lambda = dalloc<float4>(idx.size());
    sb->d_acc = dalloc<int>(3 * (size_t)n);
    sb->d_q = dalloc<float4>(n);
    sb->d_qp = dalloc<float4>(n);
    sb->d_omega = dalloc<float4>(n);
    PF_HIP(hipMemcpy(sb->d_csr_off, off.data(), (n + 1) * 4, hipMemcpyHostToDevice));
    PF_HIP(hipMemcpy(sb->d_csr_idx, idx.data(), idx.size() * 4, hipMemcpyHostToDevice));
    PF_HIP(hipMemcpy(sb->d_rest_center, rc.data(), (size_t)n * 16, hipMemcpyHostToDevice));
    // Orientations start at the identity (particles.ts:104-112).
    std::vector<float4> ident(n, make_float4(0.f, 0.f, 0.f, 1.f));
    PF_HIP(hipMemcpy(sb->d_q, ident.data(), (size_t)n * 16, hipMemcpyHostToDevice));
    PF_HIP(hipMemcpy(sb->d_qp, ident.data(), (size_t)n * 16, hipMemcpyHostToDevice));
  }
  *out = sb.release();
  PF_API_END
}

int pf_softbody_set_self_collision(pf_softbody* sb, int enabled) {
  PF_API_BEGIN
  if (!sb || sb->disposed) fail(PF_ESTATE, "SoftbodySystem: the system has been disposed");
  if (sb->loop) fail(PF_ESTATE, "SoftbodySystem: selfCollision is fixed once a SimLoop is built");
  sb->self_collision = enabled != 0;
  PF_API_END
}

int pf_softbody_set_compliance(pf_softbody* sb, uint32_t index, double compliance) {
  PF_API_BEGIN
  if (!sb || sb->disposed) fail(PF_ESTATE, "SoftbodySystem: disposed");
  if (index >= sb->bodies.size()) fail(PF_EINVAL, "SoftbodySystem: no body %u (have %zu)", index, sb->bodies.size());
  if (!(compliance >= 0) || !std::isfinite(compliance)) fail(PF_EINVAL, "SoftbodySystem: compliance must be ≥ 0, got %g", compliance);
  sb->bodies[index].compliance = compliance;
  const float c = (float)compliance;
  PF_HIP(hipStreamSynchronize(sb->ps->ctx->stream));
  PF_HIP(hipMemcpy(sb->d_compliance + index, &c, 4, hipMemcpyHostToDevice));
  const auto& bd = sb->bodies[index];
  for (uint32_t i = bd.start; i < bd.start + bd.count; ++i) sb->rest[4 * (size_t)i + 3] = c;
  PF_HIP(hipMemcpy(sb->d_rest + bd.start, sb->rest.data() + 4 * (size_t)bd.start, (size_t)bd.count * 16, hipMemcpyHostToDevice));
  PF_API_END
}

int pf_softbody_readback_frames(pf_softbody* sb, float* centers4, float* rotations12) {
  PF_API_BEGIN
  if (!sb || sb->disposed) fail(PF_ESTATE, "SoftbodySystem: disposed");
  PF_HIP(hipStreamSynchronize(sb->ps->ctx->stream));
  const size_t B = sb->bodies.size();
  if (centers4) PF_HIP(hipMemcpy(centers4, sb->d_centers, B * 16, hipMemcpyDeviceToHost));
  if (rotations12) PF_HIP(hipMemcpy(rotations12, sb->d_rotations, B * 48, hipMemcpyDeviceToHost));
  PF_API_END
}

int pf_softbody_destroy(pf_softbody* sb) {
  PF_API_BEGIN
  if (!sb || sb->disposed) return PF_OK;
  sb->disposed = true;
  dfree(sb->d_start); dfree(sb->d_count); dfree(sb->d_compliance); dfree(sb->d_rest); dfree(sb->d_weights);
  dfree(sb->d_lambda); dfree(sb->d_centers); dfree(sb->d_rotations);
  dfree(sb->d_csr_off); dfree(sb->d_csr_idx); dfree(sb->d_rest_center); dfree(sb->d_group_center);
  dfree(sb->d_group_rot); dfree(sb->d_pair_lambda); dfree(sb->d_acc); dfree(sb->d_q); dfree(sb->d_qp);
  dfree(sb->d_omega);
  PF_API_END
}

// ---------------------------------------------------------------------------
// SimLoop
// ---------------------------------------------------------------------------
static void fluid_write_params(pf_fluid* f) {
  pf_particles* ps = f->ps;
  FluidDev& F = ps->host.fluid[f->slot];
  F.start = f->start;
  F.count = f->count;
  uint32_t fx = 0;
  if (f->present & PF_FLUID_VISCOSITY) fx |= PF_FX_VISCOSITY;
  if (f->present & PF_FLUID_VORTICITY) fx |= PF_FX_VORTICITY;
  if (f->present & PF_FLUID_SURFACE_TENSION) fx |= PF_FX_SURFACE_TENSION;
  if (f->present & PF_FLUID_ADHESION) fx |= PF_FX_ADHESION;
  F.effects = fx;
  F.rho0 = (float)f->rest_density;
  F.pv = (float)(f->spacing * f->spacing * f->spacing);
  F.mass = (float)f->mass;
  F.inv_mass = 1.0f / F.mass;
  const double h = f->h;
  // kernels.ts:33-35: coefficients in double, uploaded as f32.
  F.h = (float)h;
  F.h2 = (float)(h * h);
  F.poly6 = (float)(315.0 / (64.0 * M_PI * std::pow(h, 9)));
  F.spiky = (float)(45.0 / (M_PI * std::pow(h, 6)));
  F.compliance = (float)f->compliance;
  F.gamma = (float)f->surface_tension;
  F.visc = (float)f->viscosity;
  F.vort = (float)f->vorticity;
  F.adhesion = (float)f->adhesion;
  // cohesion.ts:47, 53; adhesion.ts:17-24
  F.coh_coef = (float)(32.0 / (M_PI * std::pow(h, 9)));
  F.coh_h6_64 = (float)(std::pow(h, 6) / 64.0);
  F.adh_coef = (float)(0.007 / std::pow(h, 3.25));
  F.adh_m4_over_h = (float)(-4.0 / h);
  F.adh_2h = (float)(2.0 * h);
  // Accumulator(50) for impulses, Accumulator(10) for reactions (accumulator.ts:55-56).
  const double s50 = std::floor(1073741824.0 / 50.0), s10 = std::floor(1073741824.0 / 10.0);
  F.imp_scale = (float)s50;
  F.imp_inv = (float)(1.0 / s50);
  F.rea_scale = (float)s10;
  F.rea_inv = (float)(1.0 / s10);
  F.nboundary = (uint32_t)f->boundaries.size();
  F.bdynamic = 0;
  F.rcover_lo = 0xffffffffu;
  F.rcover_hi = 0;
  for (uint32_t b = 0; b < F.nboundary; ++b) {
    F.rcover_lo = std::min(F.rcover_lo, f->boundaries[b].start);
    F.rcover_hi = std::max(F.rcover_hi, f->boundaries[b].start + f->boundaries[b].count);
    F.bstart[b] = f->boundaries[b].start;
    F.bcount[b] = f->boundaries[b].count;
    if (f->boundaries[b].dynamic) F.bdynamic |= 1u << b;
  }
  F.nl = f->nl;
  F.nl_count = f->nl_count;
  F.nl16 = f->nl16;
  F.tu_limit = f->tu_limit;
  F.tu_size = f->tu_size;
  F.tu_slots = f->tu_slots;
  if (f->isv) {
    // rate = (nu * dt) * (10 / h^2), the constant computed in double (ViscositySolver.ts:103)
    F.isv_iters = f->isv->iterations;
    F.isv_nu = (float)f->isv->viscosity;
    F.isv_c10 = (float)(10.0 / (h * h));
    F.isv_rhs = f->isv->rhs;
    F.isv_a = f->isv->a;
    F.isv_b = f->isv->b;
  } else {
    F.isv_iters = 0;
    F.isv_rhs = F.isv_a = F.isv_b = nullptr;
  }
  F.normal = f->normal;
  F.omega = f->omega;
  F.dv = f->dv;
  F.reaction = f->reaction;
  ps->dirty = true;
}

static PlanShape loop_shape(const pf_loop* L) {
  PlanShape sh;
  sh.substeps = L->substeps;
  sh.iterations = L->iterations;
  sh.nfluids = (uint32_t)L->fluids.size();
  for (uint32_t f = 0; f < sh.nfluids; ++f) sh.fx[f] = L->ps->host.fluid[f].effects;
  sh.walls = (L->prims && !L->prims->prims.empty()) || !L->sdfs.empty();
  sh.contact = L->contact;
  sh.grid = loop_has_grid(L);
  sh.soft = L->soft != nullptr;
  sh.local_soft = L->soft != nullptr && L->soft->mode == PF_SHAPE_MATCHING_LOCAL;
  sh.any_boundary = L->ps->host.has_boundary != 0;
  for (uint32_t f = 0; f < sh.nfluids; ++f) {
    sh.nb[f] = L->ps->host.fluid[f].nboundary;
    sh.bdyn[f] = L->ps->host.fluid[f].bdynamic;
    sh.tu_cap[f] = std::min<uint32_t>(std::max<uint32_t>(L->ps->host.fluid[f].tu_limit, PF_BLOCK), PF_TU_MAX);
    sh.isv[f] = L->ps->host.fluid[f].isv_iters;
  }
  sh.grid_cap = L->grid_cap;
  return sh;
}

static void loop_set_dt(pf_loop* L, double dt) {
  SceneDev& S = L->ps->host;
  const float dtf = (float)(dt / (double)L->substeps);
  // Sweep clock per substep: frameStart sets dt*(substeps-1), every substep
  // end subtracts dt, floored at 0 (PrimitiveSet.ts:393-398).
  float clock = dtf * ((float)L->substeps - 1.0f);
  float tab[PF_MAX_SUBSTEPS];
  for (uint32_t k = 0; k < L->substeps; ++k) {
    tab[k] = clock;
    clock = std::max(clock - dtf, 0.0f);
  }
  if (S.dt != dtf || std::memcmp(S.clock, tab, sizeof(float) * L->substeps) != 0) {
    S.dt = dtf;
    std::memcpy(S.clock, tab, sizeof(float) * L->substeps);
    L->ps->dirty = true;
  }
}

int pf_loop_create(pf_particles* ps, const pf_loop_options* opt, const pf_material* mats, uint32_t nmats,
                   const pf_collider* colliders, uint32_t ncol, pf_loop** out) {
  PF_API_BEGIN
  alive(ps, "SimLoop");
  ArenaScope arena_scope(ps->arena.get());
  if (ps->loop) fail(PF_EINVAL, "SimLoop: the ParticleSystem already has a SimLoop");
  const uint32_t substeps = opt && opt->substeps ? opt->substeps : 4;
  const uint32_t iterations = opt && opt->iterations != 0xffffffffu ? opt->iterations : 2;
  if (substeps < 1 || substeps > PF_MAX_SUBSTEPS) fail(PF_EINVAL, "SimLoop: substeps must be an integer ≥ 1 (at most %d here), got %u", PF_MAX_SUBSTEPS, substeps);
  auto L = std::make_unique<pf_loop_s>();
  L->ps = ps;
  L->ctx = ps->ctx;
  L->substeps = substeps;
  L->iterations = iterations;
  if (opt) for (int a = 0; a < 3; ++a) L->gravity[a] = opt->gravity[a];
  double neighbor_radius = 0;
  for (uint32_t m = 0; m < nmats; ++m) {
    if (mats[m].kind == PF_MATERIAL_FLUID) {
      auto* f = static_cast<pf_fluid*>(mats[m].handle);
      if (!f || f->disposed) fail(PF_EINVAL, "SimLoop: a material has been disposed");
      if (f->ps != ps) fail(PF_EINVAL, "SimLoop: every material must be built for the same ParticleSystem");
      if (f->loop) fail(PF_EINVAL, "SimLoop: a fluid belongs to another SimLoop");
      if (L->soft) fail(PF_EINVAL, "SimLoop: list fluids before the soft-body system");
      if (L->fluids.size() >= PF_MAX_FLUIDS) fail(PF_EINVAL, "SimLoop: at most %d fluids per loop", PF_MAX_FLUIDS);
      L->fluids.push_back(f);
      neighbor_radius = std::max(neighbor_radius, f->h);
    } else if (mats[m].kind == PF_MATERIAL_VISCOSITY) {
      auto* v = static_cast<pf_viscosity*>(mats[m].handle);
      if (!v || v->disposed || v->fluid->disposed) fail(PF_EINVAL, "SimLoop: a material has been disposed");
      if (v->loop) fail(PF_EINVAL, "SimLoop: a ViscositySolver belongs to another SimLoop");
      if (m == 0 || mats[m - 1].kind != PF_MATERIAL_FLUID || mats[m - 1].handle != v->fluid)
        fail(PF_EINVAL, "ViscositySolver: list it directly after its FluidSystem in `materials`");
      L->viscosity.push_back(v);
    } else if (mats[m].kind == PF_MATERIAL_SOFTBODY) {
      auto* sb = static_cast<pf_softbody*>(mats[m].handle);
      if (!sb || sb->disposed) fail(PF_EINVAL, "SimLoop: a material has been disposed");
      if (sb->ps != ps) fail(PF_EINVAL, "SimLoop: every material must be built for the same ParticleSystem");
      if (L->soft) fail(PF_EINVAL, "SimLoop: one SoftbodySystem per loop");
      L->soft = sb;
    } else {
      fail(PF_EINVAL, "SimLoop: unknown material kind %u", mats[m].kind);
    }
  }
  for (uint32_t k = 0; k < ncol; ++k) {
    if (colliders[k].kind == PF_COLLIDER_PRIMITIVES) {
      auto* s = static_cast<pf_primitives*>(colliders[k].handle);
      if (!s || s->disposed) fail(PF_EINVAL, "SimLoop: a collider has been disposed");
      if (s->ps != ps) fail(PF_EINVAL, "SimLoop: every collider must be built for the same ParticleSystem");
      if (s->loop) fail(PF_EINVAL, "SimLoop: the collider belongs to another SimLoop");
      if (L->prims) fail(PF_EINVAL, "SimLoop: one PrimitiveSet per loop in this build");
      L->prims = s;
    } else if (colliders[k].kind == PF_COLLIDER_SDF) {
      auto* c = static_cast<pf_sdf_collider*>(colliders[k].handle);
      if (!c || c->disposed) fail(PF_EINVAL, "SimLoop: a collider has been disposed");
      if (c->ps != ps) fail(PF_EINVAL, "SimLoop: every collider must be built for the same ParticleSystem");
      if (c->loop) fail(PF_EINVAL, "SimLoop: the collider belongs to another SimLoop");
      if (L->sdfs.size() >= PF_MAX_SDF) fail(PF_EINVAL, "SimLoop: at most %d SDFColliders per loop", PF_MAX_SDF);
      L->sdfs.push_back(c);
    } else {
      fail(PF_EINVAL, "SimLoop: unknown collider kind %u", colliders[k].kind);
    }
  }
  // Contacts (loop.ts:148-153, 346-418).
  double ct_mus = 0.5, ct_muk = 0.4;
  uint32_t ct_max = 0;
  if (opt && opt->contact) {
    ct_mus = opt->contact_mu_s;
    ct_muk = opt->contact_mu_k;
    friction_defaults(ct_mus, ct_muk, "SimLoop");
    ct_max = opt->max_contacts ? opt->max_contacts : 8u * ps->n;
    if (opt->nquiet > PF_MAX_QUIET) fail(PF_EINVAL, "SimLoop: at most %d quiet ranges", PF_MAX_QUIET);
    if (opt->nquiet && !opt->quiet) fail(PF_EINVAL, "SimLoop: quiet ranges are missing");
    for (uint32_t q = 0; q < opt->nquiet; ++q) range_ok(ps, opt->quiet[2 * q], opt->quiet[2 * q + 1], "SimLoop");
    L->contact = true;
    neighbor_radius = std::max(neighbor_radius, 2.0 * ps->radius * 1.1);  // CONTACT_RADIUS_EXPANSION
  }
  // Collision groups handed out above the uploaded ones (loop.ts:164-168);
  // without selfCollision each soft body gets its own, unless the whole body
  // already shares a non-zero group (SoftbodySystem.ts:218-233).
  if (L->soft && !L->soft->self_collision) {
    uint32_t next = 0;
    for (uint32_t g : ps->cpu_groups) next = std::max(next, g);
    next += 1;
    const auto& bodies = L->soft->bodies;
    for (size_t b = 0; b < bodies.size(); ++b) {
      const uint32_t g0 = ps->cpu_groups[bodies[b].start];
      bool shared = true;
      for (uint32_t k = 1; k < bodies[b].count; ++k) shared = shared && ps->cpu_groups[bodies[b].start + k] == g0;
      if (!shared)
        fail(PF_EINVAL, "SoftbodySystem: body %zu has mixed collision groups; give all its particles one group, or none", b);
    }
    for (size_t b = 0; b < bodies.size(); ++b) {
      const uint32_t g0 = ps->cpu_groups[bodies[b].start];
      if (g0 != 0u) continue;
      const uint32_t group = next++;
      std::vector<uint32_t> g(bodies[b].count, group);
      for (uint32_t k = 0; k < bodies[b].count; ++k) ps->cpu_groups[bodies[b].start + k] = group;
      particles_scatter(ps, bodies[b].start, bodies[b].count, nullptr, nullptr, nullptr, &g, 8u);
    }
  }
  SceneDev& S = ps->host;
  const uint32_t n = ps->n;
  S.gx = (float)L->gravity[0]; S.gy = (float)L->gravity[1]; S.gz = (float)L->gravity[2];
  S.radius = (float)ps->radius;
  // Grid sized for the widest neighbour query (loop.ts:151-162).
  if (neighbor_radius > 0) {
    pf_grid_s& g = L->grid;
    g.ctx = ps->ctx;
    g.ps = ps;
    g.owned_by_loop = true;
    g.n = n;
    g.H = std::min<uint32_t>([&] { uint64_t p = 1; while (p < 2ull * n) p <<= 1; return (uint32_t)p; }(), 1u << 20);
    g.Hpad = (g.H + 1023) / 1024 * 1024;
    g.cell = neighbor_radius;
    if (opt && opt->has_hash_origin) for (int a = 0; a < 3; ++a) g.origin[a] = opt->hash_origin[a];
    S.H = g.H;
    S.hash_mask = g.H - 1;
    S.nscan_tiles = (g.H + PF_SCAN_TILE - 1) / PF_SCAN_TILE;
    S.cell = (float)neighbor_radius;
    S.ox = (float)g.origin[0]; S.oy = (float)g.origin[1]; S.oz = (float)g.origin[2];
    S.key = dalloc<uint32_t>(n);
    S.rank = dalloc<uint32_t>(n);
    S.counts = dalloc<uint32_t>(g.H);
    S.start = dalloc<uint32_t>((size_t)g.H + 1);
    S.scan_status = dalloc<unsigned long long>(S.nscan_tiles);
    S.scan_ctl = dalloc<uint32_t>(2);
    S.tmp_src = dalloc<uint32_t>(n);
    S.tmp_id = dalloc<uint32_t>(n);
    S.tmp_key = dalloc<uint32_t>(n);
    g.host = S;
  }
  S.nfluids = (uint32_t)L->fluids.size();
  if (S.nfluids && n > PF_FLUID_MAX_PARTICLES)
    fail(PF_EINVAL, "SimLoop: a system with fluids holds at most %u particles (neighbour-list offsets are 32-bit), got %u",
         (uint32_t)PF_FLUID_MAX_PARTICLES, n);
  for (uint32_t k = 0; k < S.nfluids; ++k) {
    pf_fluid* f = L->fluids[k];
    f->slot = (int)k;
    f->loop = L.get();
    f->nl = dalloc<uint32_t>((size_t)PF_MAX_NEIGHBORS * n);
    f->nl_count = dalloc<uint32_t>(n);
    const size_t ntiles = ((size_t)n + PF_BLOCK - 1) / PF_BLOCK;
    f->nl16 = dalloc<uint32_t>((size_t)(PF_MAX_NEIGHBORS / 2) * n);
    f->tu_size = dalloc<uint32_t>(ntiles);
    f->tu_slots = dalloc<uint32_t>(ntiles * PF_TU_MAX);
    if (f->present & PF_FLUID_SURFACE_TENSION) f->normal = dalloc<float4>(2 * (size_t)n);  // pair records
    if (f->present & PF_FLUID_VORTICITY) f->omega = dalloc<float4>(n);
    if (f->present & PF_FLUID_VISCOSITY) f->dv = dalloc<float4>(n);
    if (!f->boundaries.empty()) {
      f->reaction = dalloc<int>(3 * (size_t)n);
      S.has_boundary = 1u;
      for (auto& b : f->boundaries) L->static_boundaries = L->static_boundaries || !b.dynamic;
    }
    f->isv = nullptr;
    for (auto* v : L->viscosity)
      if (v->fluid == f) {
        v->loop = L.get();
        v->rhs = dalloc<float4>(n);
        v->a = dalloc<float4>(n);
        v->b = dalloc<float4>(n);
        f->isv = v;
      }
    fluid_write_params(f);
  }
  const double s10 = std::floor(1073741824.0 / 10.0), s50 = std::floor(1073741824.0 / 50.0);
  S.pos_scale = (float)s10;
  S.pos_inv = (float)(1.0 / s10);
  S.vel_scale = (float)s50;
  S.vel_inv = (float)(1.0 / s50);
  if (L->prims) {
    pf_primitives* s = L->prims;
    s->capacity = std::max<uint32_t>(1, s->reserved ? s->reserved : (uint32_t)s->prims.size());
    if (s->capacity > 32) fail(PF_EINVAL, "PrimitiveSet: at most 32 primitives per set");
    s->packed = dalloc<uint32_t>(s->capacity);
    s->d0 = dalloc<float4>(s->capacity);
    s->d1 = dalloc<float4>(s->capacity);
    s->vel = dalloc<float4>(s->capacity);
    s->rot = dalloc<float4>(s->capacity);
    s->spin = dalloc<float4>(s->capacity);
    s->loop = L.get();
    S.prim_cap = s->capacity;
    S.prim_packed = s->packed; S.prim_d0 = s->d0; S.prim_d1 = s->d1; S.prim_vel = s->vel; S.prim_rot = s->rot; S.prim_spin = s->spin;
    S.prim_lambda = dalloc<float>(2 * (size_t)s->capacity * n);
    S.prim_touch = dalloc<uint32_t>(n);
    s->dirty = true;
    s->upload();
  }
  S.nsdf = (uint32_t)L->sdfs.size();
  for (uint32_t k = 0; k < S.nsdf; ++k) {
    pf_sdf_collider* c = L->sdfs[k];
    c->loop = L.get();
    c->slot = (int)k;
    c->write(S.sdf[k]);
  }
  if (S.nsdf) {
    S.sdf_lambda = dalloc<float>(2 * (size_t)PF_MAX_SDF * n);
    S.sdf_touch = dalloc<uint32_t>(n);
  }
  if (L->soft) {
    pf_softbody* sb = L->soft;
    sb->loop = L.get();
    SoftDev& B = S.soft;
    B.nbodies = (uint32_t)sb->bodies.size();
    B.body_start = sb->d_start; B.body_count = sb->d_count; B.body_compliance = sb->d_compliance;
    B.rest = sb->d_rest; B.weights = sb->d_weights; B.lambda = sb->d_lambda;
    B.centers = sb->d_centers; B.rotations = sb->d_rotations;
    B.local = sb->mode == PF_SHAPE_MATCHING_LOCAL ? 1u : 0u;
    B.span_lo = sb->span_lo; B.span_hi = sb->span_hi;
    B.csr_off = sb->d_csr_off; B.csr_idx = sb->d_csr_idx; B.rest_center = sb->d_rest_center;
    B.group_center = sb->d_group_center; B.group_rot = sb->d_group_rot; B.pair_lambda = sb->d_pair_lambda;
    B.acc = sb->d_acc; B.q = sb->d_q; B.qp = sb->d_qp; B.omega = sb->d_omega;
    B.ai = (float)((ps->radius * ps->radius) / 5.0);
    B.maxc2 = (float)((5.0 * ps->radius) * (5.0 * ps->radius));
  }
  S.ct = ContactDev{};
  if (L->contact) {
    ContactDev& C = S.ct;
    C.on = 1u;
    C.max = ct_max;
    // Workgroups of the per-contact passes (grid-stride over the emitted
    // count): one record per emitting particle in flight, at most 1024.
    std::vector<uint8_t> quiet_mask(n, 0);
    for (auto* f : L->fluids) std::fill(quiet_mask.begin() + f->start, quiet_mask.begin() + f->start + f->count, 1);
    for (uint32_t q = 0; q < opt->nquiet; ++q)
      std::fill(quiet_mask.begin() + opt->quiet[2 * q], quiet_mask.begin() + opt->quiet[2 * q] + opt->quiet[2 * q + 1], 1);
    uint32_t emitters = 0;
    for (uint8_t m : quiet_mask) emitters += m ? 0u : 1u;
    const uint32_t want = std::min(std::max(emitters, 1u), ct_max);
    C.blocks = std::min<uint32_t>((want + PF_BLOCK - 1) / PF_BLOCK, 1024u);
    C.nquiet = opt->nquiet;
    for (uint32_t q = 0; q < opt->nquiet; ++q) {
      C.qstart[q] = opt->quiet[2 * q];
      C.qcount[q] = opt->quiet[2 * q + 1];
    }
    const double cr = 2.0 * ps->radius * 1.1;
    C.dist = (float)(2.0 * ps->radius);
    C.cr = (float)cr;
    C.cr2 = (float)(cr * cr);
    C.mus = (float)ct_mus;
    C.muk = (float)ct_muk;
    C.count = dalloc<uint32_t>(1);
    C.pair = dalloc<uint2>(ct_max);
    C.lam = dalloc<int2>(ct_max);
    C.nrm = dalloc<float4>(ct_max);
    C.acc = dalloc<int4>(n);
    PF_HIP(hipMemset(C.count, 0, sizeof(uint32_t)));
    PF_HIP(hipMemset(C.acc, 0, (size_t)n * sizeof(int4)));
    PF_HIP(hipMemset(C.nrm, 0, (size_t)ct_max * sizeof(float4)));
    PF_HIP(hipMemset(C.lam, 0, (size_t)ct_max * sizeof(int2)));
    L->nctiles = C.blocks;
    std::vector<Tile> ctl(L->nctiles);
    for (uint32_t i = 0; i < L->nctiles; ++i) ctl[i] = Tile{0u, i * PF_BLOCK};
    L->ctiles = dalloc<Tile>(L->nctiles);
    PF_HIP(hipMemcpy(L->ctiles, ctl.data(), ctl.size() * sizeof(Tile), hipMemcpyHostToDevice));
  }
  // The scatter accumulator also carries surface tension's scattered pairs.
  bool any_st = false;
  for (auto* f : L->fluids) any_st = any_st || (f->present & PF_FLUID_SURFACE_TENSION);
  if (any_st && !S.ct.acc) {
    S.ct.acc = dalloc<int4>(n);
    PF_HIP(hipMemset(S.ct.acc, 0, (size_t)n * sizeof(int4)));
  }
  // Work tiles.
  L->ntiles = (n + PF_BLOCK - 1) / PF_BLOCK;
  std::vector<Tile> tl(L->ntiles);
  for (uint32_t i = 0; i < L->ntiles; ++i) tl[i] = Tile{0u, i * PF_BLOCK};
  L->tiles = dalloc<Tile>(L->ntiles);
  PF_HIP(hipMemcpy(L->tiles, tl.data(), tl.size() * sizeof(Tile), hipMemcpyHostToDevice));
  L->nstiles = S.nscan_tiles;
  if (L->nstiles) {
    std::vector<uint2> stl(L->nstiles);
    for (uint32_t i = 0; i < L->nstiles; ++i) stl[i] = make_uint2(0u, i);
    L->stiles = dalloc<uint2>(L->nstiles);
    PF_HIP(hipMemcpy(L->stiles, stl.data(), stl.size() * sizeof(uint2), hipMemcpyHostToDevice));
  }
  if (L->soft && S.soft.local) {
    const uint32_t span = S.soft.span_hi - S.soft.span_lo;
    L->nltiles = (span + PF_BLOCK - 1) / PF_BLOCK;
    std::vector<uint2> ltl(L->nltiles);
    for (uint32_t i = 0; i < L->nltiles; ++i) ltl[i] = make_uint2(0u, S.soft.span_lo + i * PF_BLOCK);
    L->ltiles = dalloc<uint2>(L->nltiles);
    PF_HIP(hipMemcpy(L->ltiles, ltl.data(), ltl.size() * sizeof(uint2), hipMemcpyHostToDevice));
  }
  L->nbtiles = S.soft.nbodies;
  if (L->nbtiles) {
    std::vector<uint2> btl(L->nbtiles);
    for (uint32_t i = 0; i < L->nbtiles; ++i) btl[i] = make_uint2(0u, i);
    L->btiles = dalloc<uint2>(L->nbtiles);
    PF_HIP(hipMemcpy(L->btiles, btl.data(), btl.size() * sizeof(uint2), hipMemcpyHostToDevice));
  }
  ps->loop = L.get();
  ps->dirty = true;
  ps->upload_desc();
  *out = L.release();
  PF_API_END
}

static void loop_alive(pf_loop* L, const char* what) {
  if (!L) fail(PF_EINVAL, "%s: null handle", what);
  if (L->disposed) fail(PF_ESTATE, "SimLoop: the loop has been disposed");
  if (L->ps->disposed) fail(PF_ESTATE, "SimLoop: the ParticleSystem has been disposed");
  if (L->ps->server) fail(PF_ESTATE, "%s: the scene is being served; remove it from the server first", what);
}

// A loop steps only while every material and collider it was built with is
// alive: their device buffers go with them.
static void loop_members_alive(pf_loop* L) {
  for (auto* f : L->fluids)
    if (f->disposed) fail(PF_ESTATE, "SimLoop: a material has been disposed");
  if (L->soft && L->soft->disposed) fail(PF_ESTATE, "SimLoop: a material has been disposed");
  for (auto* v : L->viscosity)
    if (v->disposed) fail(PF_ESTATE, "SimLoop: a material has been disposed");
  if (L->prims && L->prims->disposed) fail(PF_ESTATE, "SimLoop: a collider has been disposed");
  for (auto* c : L->sdfs)
    if (c->disposed) fail(PF_ESTATE, "SimLoop: a collider has been disposed");
}

// Every graph of the loop is dropped; a replay still running on the stream
// finishes first, since an executable graph must not be destroyed in flight.
static void loop_invalidate_graphs(pf_loop* L) {
  bool any = false;
  for (int v = 0; v < 4; ++v) any = any || L->graph[v];
  if (any) PF_HIP(hipStreamSynchronize(L->ctx->stream));
  for (int v = 0; v < 4; ++v) {
    if (L->graph[v]) (void)hipGraphExecDestroy(L->graph[v]);
    L->graph[v] = nullptr;
    L->graph_valid[v] = false;
  }
}

// First step of a loop with static boundaries (loop.ts:236, 276-277): sort
// the committed positions (a prediction over dt = 0 leaves them in place)
// and compute the static boundary volumes from them, once.
static void loop_static_boundaries(pf_loop* L) {
  pf_particles* ps = L->ps;
  Context* ctx = L->ctx;
  hipStream_t st = ctx->stream;
  SceneDev& S = ps->host;
  const float dt_keep = S.dt;
  S.dt = 0.0f;
  ps->dirty = true;
  ps->upload_desc();
  const SceneDev* d = ps->cur();
  launch_hash(st, d, L->tiles, L->ntiles, 0, 0, 0);
  launch_scan(st, d, L->stiles, L->nstiles, ctx->scan_ctl);
  launch_place(st, d, L->tiles, L->ntiles, 0);
  launch_rank(st, d, L->tiles, L->ntiles, 0, 0, 1);
  for (uint32_t f = 0; f < (uint32_t)L->fluids.size(); ++f)
    if (S.fluid[f].nboundary) launch_boundary_volume(st, d, L->tiles, L->ntiles, (int)f, 1, 1, 0);
  PF_HIP(hipGetLastError());
  PF_HIP(hipStreamSynchronize(st));
  // The sorted state is the loop's state from here: swap both pairs.
  ps->variant ^= 3;
  S.dt = dt_keep;
  ps->dirty = true;
}

}  // extern "C"

namespace pf {
PlanShape loop_plan_shape(const pf_loop_s* L) { return loop_shape(L); }

void plan_merge(PlanShape& sh, const PlanShape& ls) {
  for (uint32_t f = 0; f < ls.nfluids; ++f) {
    sh.tu_cap[f] = f < sh.nfluids ? std::max(sh.tu_cap[f], ls.tu_cap[f]) : ls.tu_cap[f];
    sh.isv[f] = std::max(sh.isv[f], ls.isv[f]);
  }
  sh.nfluids = std::max(sh.nfluids, ls.nfluids);
  for (uint32_t f = 0; f < PF_MAX_FLUIDS; ++f) {
    sh.fx[f] |= ls.fx[f];
    sh.nb[f] = std::max(sh.nb[f], ls.nb[f]);
    sh.bdyn[f] |= ls.bdyn[f];
  }
  sh.walls = sh.walls || ls.walls;
  sh.contact = sh.contact || ls.contact;
  sh.grid = sh.grid || ls.grid;
  sh.soft = sh.soft || ls.soft;
  sh.local_soft = sh.local_soft || ls.local_soft;
  sh.any_boundary = sh.any_boundary || ls.any_boundary;
}

bool loop_has_grid(const pf_loop_s* L) { return !L->fluids.empty() || L->contact; }

void loop_append_tiles(const pf_loop_s* L, uint32_t scene, std::vector<Tile>& tl, std::vector<uint2>& stl,
                       std::vector<uint2>& btl, std::vector<uint2>& ltl, std::vector<Tile>& ctl) {
  for (uint32_t i = 0; i < L->ntiles; ++i) tl.push_back(Tile{scene, i * PF_BLOCK});
  if (L->contact)
    for (uint32_t i = 0; i < L->ps->host.ct.blocks; ++i) ctl.push_back(Tile{scene, i * PF_BLOCK});
  for (uint32_t i = 0; i < L->nstiles; ++i) stl.push_back(make_uint2(scene, i));
  for (uint32_t i = 0; i < L->nbtiles; ++i) btl.push_back(make_uint2(scene, i));
  if (L->soft && L->soft->mode == PF_SHAPE_MATCHING_LOCAL) {
    const SoftDev& SB = L->ps->host.soft;
    const uint32_t nl = (SB.span_hi - SB.span_lo + PF_BLOCK - 1) / PF_BLOCK;
    for (uint32_t i = 0; i < nl; ++i) ltl.push_back(make_uint2(scene, SB.span_lo + i * PF_BLOCK));
  }
}

// Host work before a loop's step runs inside a batch or a server launch.
// Returns true when the scene's descriptor changed.
bool loop_prepare_step(pf_loop_s* L, double dt) {
  loop_members_alive(L);
  if (L->prims) L->prims->upload();
  loop_set_dt(L, dt);
  if (L->static_boundaries && L->steps == 0) loop_static_boundaries(L);
  return L->ps->dirty;
}

// Host bookkeeping after a loop's step was enqueued in a batch or a server launch.
void loop_finish_step(pf_loop_s* L, int flips) {
  L->ps->variant ^= flips;
  L->ps->stepped = true;
  L->last_flag_half = (L->substeps - 1) & 1u;
  L->steps += 1;
}
}  // namespace pf

extern "C" {

int pf_loop_step(pf_loop* L, double dt) {
  PF_API_BEGIN
  loop_alive(L, "SimLoop.step");
  loop_members_alive(L);
  if (!(dt > 0) || !std::isfinite(dt)) fail(PF_EINVAL, "SimLoop: step dt must be a positive finite number, got %g", dt);
  pf_particles* ps = L->ps;
  Context* ctx = L->ctx;
  if (L->prims) {
    const uint32_t before = ps->host.nprim;
    L->prims->upload();
    if (ps->host.nprim != before) loop_invalidate_graphs(L);
  }
  loop_set_dt(L, dt);
  if (L->static_boundaries && L->steps == 0) loop_static_boundaries(L);
  ps->upload_desc();
  const PlanShape sh = loop_shape(L);
  const int v = ps->variant;
  if (L->use_graph) {
    if (!L->graph_valid[v]) {
      hipGraph_t g;
      PF_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
      ctx->capturing = true;
      try {
        L->kernels_per_step = enqueue_step(ctx, ps->cur(), sh, L->tiles, L->ntiles, L->stiles, L->nstiles,
                                           L->btiles, L->nbtiles, L->ltiles, L->nltiles, L->ctiles, L->nctiles,
                                           ctx->scan_ctl);
      } catch (...) {
        ctx->capturing = false;
        hipGraph_t dummy;
        (void)hipStreamEndCapture(ctx->stream, &dummy);
        throw;
      }
      ctx->capturing = false;
      PF_HIP(hipStreamEndCapture(ctx->stream, &g));
      PF_HIP(hipGraphInstantiate(&L->graph[v], g, nullptr, nullptr, 0));
      PF_HIP(hipGraphDestroy(g));
      L->graph_valid[v] = true;
    }
    PF_HIP(hipGraphLaunch(L->graph[v], ctx->stream));
  } else {
    L->kernels_per_step = enqueue_step(ctx, ps->cur(), sh, L->tiles, L->ntiles, L->stiles, L->nstiles, L->btiles,
                                       L->nbtiles, L->ltiles, L->nltiles, L->ctiles, L->nctiles, ctx->scan_ctl);
  }
  PF_HIP(hipGetLastError());
  ps->variant ^= step_variant_flips(sh);
  ps->stepped = true;
  L->last_flag_half = (L->substeps - 1) & 1u;
  L->steps += 1;
  PF_API_END
}

int pf_loop_synchronize(pf_loop* L) {
  PF_API_BEGIN
  loop_alive(L, "SimLoop");
  PF_HIP(hipStreamSynchronize(L->ctx->stream));
  PF_API_END
}

// Completion fences: an event recorded on the loop's stream behind the work
// enqueued so far. Waiting on one touches neither the loop nor its stream,
// so a host thread may wait while the loop's owner enqueues (or captures)
// more work.
static pf_fence* fence_on(Context* ctx) {
  if (ctx->capturing) fail(PF_ESTATE, "SimLoop: no fence while a step is being captured");
  hipEvent_t ev = nullptr;
  PF_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
  const hipError_t e = hipEventRecord(ev, ctx->stream);
  if (e != hipSuccess) {
    (void)hipEventDestroy(ev);
    PF_HIP(e);
  }
  auto* f = new pf_fence_s();
  f->ev = ev;
  return f;
}

int pf_loop_fence(pf_loop* L, pf_fence** out) {
  PF_API_BEGIN
  loop_alive(L, "SimLoop");
  if (!out) fail(PF_EINVAL, "SimLoop: null handle");
  *out = fence_on(L->ctx);
  PF_API_END
}

int pf_context_fence(pf_context* ctx, pf_fence** out) {
  PF_API_BEGIN
  if (!ctx || !out) fail(PF_EINVAL, "pf_context_fence: null handle");
  *out = fence_on(ctx);
  PF_API_END
}

int pf_fence_wait(pf_fence* f) {
  PF_API_BEGIN
  if (!f) fail(PF_EINVAL, "fence: null handle");
  PF_HIP(hipEventSynchronize(f->ev));
  PF_API_END
}

int pf_fence_query(pf_fence* f, int* done) {
  PF_API_BEGIN
  if (!f || !done) fail(PF_EINVAL, "fence: null handle");
  const hipError_t e = hipEventQuery(f->ev);
  if (e != hipSuccess && e != hipErrorNotReady) PF_HIP(e);
  *done = e == hipSuccess ? 1 : 0;
  PF_API_END
}

int pf_fence_destroy(pf_fence* f) {
  PF_API_BEGIN
  if (!f) return PF_OK;
  (void)hipEventDestroy(f->ev);
  delete f;
  PF_API_END
}

int pf_loop_readback_overflow(pf_loop* L, pf_overflow* o) {
  PF_API_BEGIN
  loop_alive(L, "SimLoop.readbackOverflow");
  PF_HIP(hipStreamSynchronize(L->ctx->stream));
  uint32_t fl[2 * PF_NUM_FLAGS];
  PF_HIP(hipMemcpy(fl, L->ps->host.flags, sizeof fl, hipMemcpyDeviceToHost));
  const uint32_t* last = fl + L->last_flag_half * PF_NUM_FLAGS;
  std::memset(o, 0, sizeof *o);
  o->positions = fl[PF_FLAG_POSITIONS] != 0;
  o->velocities = fl[PF_FLAG_VELOCITIES] != 0;
  o->grid = last[PF_FLAG_GRID] != 0;
  o->contacts = 0;
  if (L->contact) {  // the last substep found more pairs than it could keep (loop.ts:292)
    uint32_t cnt = 0;
    PF_HIP(hipMemcpy(&cnt, L->ps->host.ct.count, sizeof cnt, hipMemcpyDeviceToHost));
    o->contacts = cnt > L->ps->host.ct.max;
  }
  o->neighbors = last[PF_FLAG_NEIGHBORS] != 0;
  o->impulses = fl[PF_FLAG_IMPULSES] != 0;
  o->long_bucket = last[PF_FLAG_LONG_BUCKET] != 0;
  PF_API_END
}

int pf_loop_readback_contacts(pf_loop* L, uint32_t* count, uint32_t* pairs, uint32_t cap, uint32_t* stored) {
  PF_API_BEGIN
  loop_alive(L, "SimLoop.contacts");
  if (!L->contact) fail(PF_EINVAL, "SimLoop: the loop has no contacts");
  PF_HIP(hipStreamSynchronize(L->ctx->stream));
  const ContactDev& C = L->ps->host.ct;
  uint32_t cnt = 0;
  PF_HIP(hipMemcpy(&cnt, C.count, sizeof cnt, hipMemcpyDeviceToHost));
  const uint32_t m = pairs ? std::min(std::min(cnt, C.max), cap) : 0u;
  if (m) {
    std::vector<uint2> pr(m);
    PF_HIP(hipMemcpy(pr.data(), C.pair, m * sizeof(uint2), hipMemcpyDeviceToHost));
    // Records hold slots of the last substep's order, which the state keeps
    // until the next step sorts again.
    std::vector<uint32_t> ids(L->ps->n);
    const SceneDev d = L->ps->desc(L->ps->variant);
    PF_HIP(hipMemcpy(ids.data(), d.id[0], ids.size() * sizeof(uint32_t), hipMemcpyDeviceToHost));
    for (uint32_t k = 0; k < m; ++k) {
      pairs[2 * k] = ids[pr[k].x];
      pairs[2 * k + 1] = ids[pr[k].y];
    }
  }
  if (count) *count = cnt;
  if (stored) *stored = m;
  PF_API_END
}

int pf_loop_set_substeps(pf_loop* L, uint32_t substeps) {
  PF_API_BEGIN
  loop_alive(L, "SimLoop.substeps");
  if (substeps < 1 || substeps > PF_MAX_SUBSTEPS) fail(PF_EINVAL, "SimLoop: substeps must be an integer ≥ 1, got %u", substeps);
  if (substeps != L->substeps) {
    L->substeps = substeps;
    L->last_dt = -1;
    loop_invalidate_graphs(L);
  }
  PF_API_END
}

int pf_loop_set_gravity(pf_loop* L, const double g[3]) {
  PF_API_BEGIN
  loop_alive(L, "SimLoop.gravity");
  for (int a = 0; a < 3; ++a) L->gravity[a] = g[a];
  SceneDev& S = L->ps->host;
  const float gx = (float)g[0], gy = (float)g[1], gz = (float)g[2];
  if (S.gx != gx || S.gy != gy || S.gz != gz) {
    S.gx = gx; S.gy = gy; S.gz = gz;
    L->ps->dirty = true;
  }
  PF_API_END
}

int pf_loop_info(pf_loop* L, uint32_t* substeps, uint32_t* iterations, double* cell, uint32_t* H,
                 uint32_t* kernels_per_step) {
  PF_API_BEGIN
  loop_alive(L, "SimLoop");
  if (substeps) *substeps = L->substeps;
  if (iterations) *iterations = L->iterations;
  if (cell) *cell = L->grid.cell;
  if (H) *H = L->grid.H;
  if (kernels_per_step) *kernels_per_step = L->kernels_per_step;
  PF_API_END
}

int pf_loop_grid(pf_loop* L, pf_grid** out) {
  PF_API_BEGIN
  loop_alive(L, "SimLoop.hashGrid");
  if (!L->grid.owned_by_loop) fail(PF_EINVAL, "SimLoop: no material needs a neighbour grid");
  L->grid.host = L->ps->host;
  *out = &L->grid;
  PF_API_END
}

int pf_loop_set_graph(pf_loop* L, int enabled) {
  PF_API_BEGIN
  loop_alive(L, "SimLoop");
  if ((enabled != 0) != L->use_graph) loop_invalidate_graphs(L);
  L->use_graph = enabled != 0;
  PF_API_END
}

int pf_loop_set_launch_shape(pf_loop* L, uint32_t workgroups) {
  PF_API_BEGIN
  loop_alive(L, "SimLoop");
  check_launch_shape(workgroups, "SimLoop");
  if (workgroups != L->grid_cap) loop_invalidate_graphs(L);
  L->grid_cap = workgroups;
  PF_API_END
}

int pf_loop_destroy(pf_loop* L) {
  PF_API_BEGIN
  if (!L || L->disposed) return PF_OK;
  (void)hipStreamSynchronize(L->ctx->stream);
  L->disposed = true;
  loop_invalidate_graphs(L);
  SceneDev& S = L->ps->host;
  dfree(S.key); dfree(S.rank); dfree(S.counts); dfree(S.start); dfree(S.scan_status); dfree(S.scan_ctl);
  dfree(S.tmp_src); dfree(S.tmp_id); dfree(S.tmp_key); dfree(S.prim_lambda); dfree(S.prim_touch);
  S.key = S.rank = S.counts = S.start = S.tmp_src = S.tmp_id = S.tmp_key = nullptr;
  S.scan_status = nullptr; S.scan_ctl = nullptr; S.prim_lambda = nullptr; S.prim_touch = nullptr;
  dfree(L->tiles); dfree(L->stiles); dfree(L->btiles); dfree(L->ltiles); dfree(L->ctiles);
  dfree(S.ct.count); dfree(S.ct.pair); dfree(S.ct.lam); dfree(S.ct.nrm); dfree(S.ct.acc);
  S.ct = ContactDev{};
  for (auto* f : L->fluids) {
    f->loop = nullptr;
    dfree(f->nl); dfree(f->nl_count); dfree(f->normal); dfree(f->omega); dfree(f->dv);
    dfree(f->nl16); dfree(f->tu_size); dfree(f->tu_slots);
    f->nl16 = f->tu_size = f->tu_slots = nullptr;
    f->nl = nullptr; f->nl_count = nullptr; f->normal = f->omega = f->dv = nullptr;
  }
  if (L->prims) L->prims->loop = nullptr;
  for (auto* c : L->sdfs) {
    c->loop = nullptr;
    c->slot = -1;
  }
  dfree(S.sdf_lambda); dfree(S.sdf_touch);
  S.sdf_lambda = nullptr; S.sdf_touch = nullptr; S.nsdf = 0;
  for (auto* f : L->fluids) { dfree(f->reaction); f->reaction = nullptr; f->isv = nullptr; }
  for (auto* v : L->viscosity) {
    dfree(v->rhs); dfree(v->a); dfree(v->b);
    v->rhs = v->a = v->b = nullptr;
    v->loop = nullptr;
  }
  if (L->soft) L->soft->loop = nullptr;
  L->ps->loop = nullptr;
  L->ps->dirty = true;
  PF_API_END
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Batched form
// ---------------------------------------------------------------------------

// Scenes whose substep schedules are the same (union capacity and launch
// shape aside) share a batch group.
static bool same_schedule(const PlanShape& a, const PlanShape& b) {
  if (a.nfluids != b.nfluids || a.any_boundary != b.any_boundary || a.local_soft != b.local_soft ||
      a.walls != b.walls || a.soft != b.soft || a.grid != b.grid || a.contact != b.contact)
    return false;
  for (uint32_t f = 0; f < PF_MAX_FLUIDS; ++f)
    if (a.fx[f] != b.fx[f] || a.nb[f] != b.nb[f] || a.bdyn[f] != b.bdyn[f] || a.isv[f] != b.isv[f]) return false;
  return true;
}

template <typename T>
static T* upload_list(const std::vector<T>& v, uint32_t* count) {
  *count = (uint32_t)v.size();
  if (v.empty()) return nullptr;
  T* d = dalloc<T>(v.size());
  PF_HIP(hipMemcpy(d, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
  return d;
}

static void batch_group_free(BatchGroup& g) {
  dfree(g.tiles); dfree(g.stiles); dfree(g.btiles); dfree(g.ltiles); dfree(g.ctiles);
  g = BatchGroup{};
}

extern "C" {

int pf_batch_create(pf_context* ctx, pf_loop* const* loops, uint32_t n, pf_batch** out) {
  PF_API_BEGIN
  if (!ctx || !loops || n == 0) fail(PF_EINVAL, "Batch: at least one loop is required");
  auto B = std::unique_ptr<pf_batch_s, void (*)(pf_batch_s*)>(new pf_batch_s(), [](pf_batch_s* b) {
    for (auto& g : b->groups) batch_group_free(g);
    dfree(b->dev);
    dfree(b->scan_ctl);
    delete b;
  });
  B->ctx = ctx;
  PlanShape sh;
  sh.substeps = loops[0]->substeps;
  sh.iterations = loops[0]->iterations;
  std::vector<PlanShape> keys;  // each group's first scene's shape
  for (uint32_t s = 0; s < n; ++s) {
    pf_loop* L = loops[s];
    loop_alive(L, "Batch");
    if (L->ctx != ctx) fail(PF_EINVAL, "Batch: every loop must use the batch's context");
    if (L->substeps != sh.substeps || L->iterations != sh.iterations)
      fail(PF_EINVAL, "Batch: every loop must share substeps and iterations");
    for (uint32_t t = 0; t < s; ++t)
      if (loops[t] == L) fail(PF_EINVAL, "Batch: a loop appears twice");
    const PlanShape own = loop_shape(L);
    plan_merge(sh, own);
    size_t g = 0;
    while (g < keys.size() && !same_schedule(keys[g], own)) ++g;
    if (g == keys.size()) {
      keys.push_back(own);
      BatchGroup grp;
      grp.shape.substeps = sh.substeps;
      grp.shape.iterations = sh.iterations;
      B->groups.push_back(grp);
    }
    B->groups[g].scenes.push_back(s);
    plan_merge(B->groups[g].shape, own);
    B->loops.push_back(L);
  }
  for (auto* L : B->loops)
    if (loop_has_grid(L) != sh.grid)
      fail(PF_EINVAL, "Batch: mixing scenes with and without a neighbour grid (fluids or contacts) is not supported");
  B->shape = sh;
  // Each group adds its own launches: many different schedules run as one
  // group under the merged schedule instead.
  constexpr size_t kMaxGroups = 4;
  if (B->groups.size() > kMaxGroups) {
    BatchGroup all;
    all.shape = sh;
    for (uint32_t s = 0; s < n; ++s) all.scenes.push_back(s);
    B->groups.assign(1, all);
  }
  for (auto& g : B->groups) {
    std::vector<Tile> tl, ctl;
    std::vector<uint2> stl, btl, ltl;
    for (uint32_t s : g.scenes) loop_append_tiles(B->loops[s], s, tl, stl, btl, ltl, ctl);
    g.tiles = upload_list(tl, &g.ntiles);
    g.stiles = upload_list(stl, &g.nstiles);
    g.btiles = upload_list(btl, &g.nbtiles);
    g.ltiles = upload_list(ltl, &g.nltiles);
    g.ctiles = upload_list(ctl, &g.nctiles);
  }
  B->dev = dalloc<SceneDev>(n);
  B->scan_ctl = dalloc<uint32_t>(4);
  *out = B.release();
  PF_API_END
}

int pf_batch_step(pf_batch* B, double dt) {
  PF_API_BEGIN
  if (!B) fail(PF_EINVAL, "Batch: null handle");
  if (!(dt > 0) || !std::isfinite(dt)) fail(PF_EINVAL, "Batch: step dt must be a positive finite number, got %g", dt);
  Context* ctx = B->ctx;
  bool changed = false;
  for (auto* L : B->loops) {
    loop_alive(L, "Batch");
    changed = loop_prepare_step(L, dt) || changed;
  }
  // A changed tile-union limit changes the launches' LDS capacity.
  bool cap_changed = false;
  for (auto& g : B->groups) {
    PlanShape now;
    for (uint32_t s : g.scenes) plan_merge(now, loop_shape(B->loops[s]));
    if (std::memcmp(now.tu_cap, g.shape.tu_cap, sizeof now.tu_cap) != 0) {
      std::memcpy(g.shape.tu_cap, now.tu_cap, sizeof now.tu_cap);
      cap_changed = true;
    }
  }
  if (cap_changed) {
    if (B->graph) {
      PF_HIP(hipStreamSynchronize(ctx->stream));  // not destroyed while a replay runs
      (void)hipGraphExecDestroy(B->graph);
      B->graph = nullptr;
    }
    B->graph_valid = false;
  }
  if (changed || B->desc_dirty) {
    PF_HIP(hipStreamSynchronize(ctx->stream));
    std::vector<SceneDev> descs;
    for (auto* L : B->loops) {
      L->ps->upload_desc();
      descs.push_back(L->ps->desc(L->ps->variant));
    }
    PF_HIP(hipMemcpy(B->dev, descs.data(), descs.size() * sizeof(SceneDev), hipMemcpyHostToDevice));
    B->desc_dirty = false;
  }
  // The groups' steps one after the other on the stream (one graph).
  auto enqueue_groups = [&] {
    for (const auto& g : B->groups) {
      PlanShape sh = g.shape;
      sh.grid_cap = B->shape.grid_cap;
      enqueue_step(ctx, B->dev, sh, g.tiles, g.ntiles, g.stiles, g.nstiles, g.btiles, g.nbtiles, g.ltiles,
                   g.nltiles, g.ctiles, g.nctiles, B->scan_ctl);
    }
  };
  if (B->use_graph) {
    if (!B->graph_valid) {
      hipGraph_t g;
      PF_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
      ctx->capturing = true;
      try {
        enqueue_groups();
      } catch (...) {
        ctx->capturing = false;
        hipGraph_t dummy;
        (void)hipStreamEndCapture(ctx->stream, &dummy);
        throw;
      }
      ctx->capturing = false;
      PF_HIP(hipStreamEndCapture(ctx->stream, &g));
      PF_HIP(hipGraphInstantiate(&B->graph, g, nullptr, nullptr, 0));
      PF_HIP(hipGraphDestroy(g));
      B->graph_valid = true;
    }
    PF_HIP(hipGraphLaunch(B->graph, ctx->stream));
  } else {
    enqueue_groups();
  }
  PF_HIP(hipGetLastError());
  for (const auto& g : B->groups) {
    const int flips = step_variant_flips(g.shape);
    for (uint32_t s : g.scenes) loop_finish_step(B->loops[s], flips);
    // The graph reads the descriptor array at run time; swapped buffer pairs
    // only need the array re-uploaded before the next step.
    if (flips) B->desc_dirty = true;
  }
  PF_API_END
}

// ---------------------------------------------------------------------------
// Autotuner: the tile-union LDS capacity (PF_BLOCK..PF_TU_MAX entries per
// 256-slot tile) trades occupancy of the neighbour kernels against tiles that
// fall back to global gathers. Each candidate is timed over `steps` steps of
// the real workload (after one untimed step that captures its graph), in a
// forward then a backward sweep so the scenes' drift during tuning averages
// out; the fastest is kept. The tuning steps advance the scenes.
// ---------------------------------------------------------------------------
}  // extern "C"

static void tune_set_limit(const std::vector<pf_loop*>& loops, uint32_t cap) {
  for (auto* L : loops)
    for (auto* f : L->fluids)
      if (pf_fluid_set_tile_union_limit(f, cap) != PF_OK) throw PfError{PF_EINVAL};
}

// Times each candidate setting over `steps` steps (after one untimed step
// that captures its graph), in a forward then a backward sweep, and keeps the
// fastest: set(candidate) applies a setting, step() enqueues one step.
template <typename SetFn, typename StepFn>
static uint32_t tune_run(Context* ctx, const uint32_t* cand, uint32_t n, uint32_t steps, float* ms_out, SetFn set,
                         StepFn step) {
  if (steps == 0) steps = 1;
  std::vector<double> total(n, 0.0);
  hipEvent_t e0, e1;
  PF_HIP(hipEventCreate(&e0));
  PF_HIP(hipEventCreate(&e1));
  try {
    for (int pass = 0; pass < 2; ++pass) {
      for (uint32_t k = 0; k < n; ++k) {
        const uint32_t i = pass == 0 ? k : n - 1 - k;
        set(cand[i]);
        step();  // captures the candidate's graph
        PF_HIP(hipEventRecord(e0, ctx->stream));
        for (uint32_t t = 0; t < steps; ++t) step();
        PF_HIP(hipEventRecord(e1, ctx->stream));
        PF_HIP(hipEventSynchronize(e1));
        float ms = 0.f;
        PF_HIP(hipEventElapsedTime(&ms, e0, e1));
        total[i] += ms / steps;
      }
    }
  } catch (...) {
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    throw;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  uint32_t best = 0;
  for (uint32_t i = 0; i < n; ++i) {
    if (ms_out) ms_out[i] = (float)(total[i] / 2.0);
    if (total[i] < total[best]) best = i;
  }
  set(cand[best]);
  return cand[best];
}

static void tune_check_limits(const uint32_t* cand, uint32_t n) {
  if (!cand || n == 0) fail(PF_EINVAL, "Autotune: at least one candidate is required");
  for (uint32_t i = 0; i < n; ++i)
    if (cand[i] < PF_BLOCK || cand[i] > PF_TU_MAX)
      fail(PF_EINVAL, "Autotune: tile-union capacity %u outside [%u, %u]", cand[i], PF_BLOCK, PF_TU_MAX);
}

static void tune_check_shapes(const uint32_t* cand, uint32_t n) {
  if (!cand || n == 0) fail(PF_EINVAL, "Autotune: at least one candidate is required");
  for (uint32_t i = 0; i < n; ++i) check_launch_shape(cand[i], "Autotune");
}

extern "C" {

int pf_loop_autotune(pf_loop* L, double dt, const uint32_t* candidates, uint32_t n, uint32_t steps,
                     uint32_t* best, float* ms_per_step) {
  PF_API_BEGIN
  loop_alive(L, "SimLoop.autotune");
  if (L->fluids.empty()) fail(PF_EINVAL, "SimLoop.autotune: the loop has no fluid to tune");
  tune_check_limits(candidates, n);
  std::vector<pf_loop*> loops{L};
  const uint32_t b = tune_run(
      L->ctx, candidates, n, steps, ms_per_step, [&](uint32_t c) { tune_set_limit(loops, c); },
      [&] {
        if (pf_loop_step(L, dt) != PF_OK) throw PfError{PF_EINVAL};
      });
  if (best) *best = b;
  PF_API_END
}

int pf_batch_autotune(pf_batch* B, double dt, const uint32_t* candidates, uint32_t n, uint32_t steps,
                      uint32_t* best, float* ms_per_step) {
  PF_API_BEGIN
  if (!B) fail(PF_EINVAL, "Batch: null handle");
  if (B->shape.nfluids == 0) fail(PF_EINVAL, "Batch.autotune: the batch has no fluid to tune");
  tune_check_limits(candidates, n);
  const uint32_t b = tune_run(
      B->ctx, candidates, n, steps, ms_per_step, [&](uint32_t c) { tune_set_limit(B->loops, c); },
      [&] {
        if (pf_batch_step(B, dt) != PF_OK) throw PfError{PF_EINVAL};
      });
  if (best) *best = b;
  PF_API_END
}

int pf_loop_autotune_launch(pf_loop* L, double dt, const uint32_t* candidates, uint32_t n, uint32_t steps,
                            uint32_t* best, float* ms_per_step) {
  PF_API_BEGIN
  loop_alive(L, "SimLoop.autotuneLaunch");
  tune_check_shapes(candidates, n);
  const uint32_t b = tune_run(
      L->ctx, candidates, n, steps, ms_per_step,
      [&](uint32_t w) {
        if (pf_loop_set_launch_shape(L, w) != PF_OK) throw PfError{PF_EINVAL};
      },
      [&] {
        if (pf_loop_step(L, dt) != PF_OK) throw PfError{PF_EINVAL};
      });
  if (best) *best = b;
  PF_API_END
}

int pf_batch_autotune_launch(pf_batch* B, double dt, const uint32_t* candidates, uint32_t n, uint32_t steps,
                             uint32_t* best, float* ms_per_step) {
  PF_API_BEGIN
  if (!B) fail(PF_EINVAL, "Batch: null handle");
  tune_check_shapes(candidates, n);
  const uint32_t b = tune_run(
      B->ctx, candidates, n, steps, ms_per_step,
      [&](uint32_t w) {
        if (pf_batch_set_launch_shape(B, w) != PF_OK) throw PfError{PF_EINVAL};
      },
      [&] {
        if (pf_batch_step(B, dt) != PF_OK) throw PfError{PF_EINVAL};
      });
  if (best) *best = b;
  PF_API_END
}

int pf_batch_set_launch_shape(pf_batch* B, uint32_t workgroups) {
  PF_API_BEGIN
  if (!B) fail(PF_EINVAL, "Batch: null handle");
  check_launch_shape(workgroups, "Batch");
  if (workgroups != B->shape.grid_cap && B->graph) {
    PF_HIP(hipStreamSynchronize(B->ctx->stream));  // not destroyed while a replay runs
    (void)hipGraphExecDestroy(B->graph);
    B->graph = nullptr;
  }
  if (workgroups != B->shape.grid_cap) B->graph_valid = false;
  B->shape.grid_cap = workgroups;
  PF_API_END
}

int pf_batch_set_graph(pf_batch* B, int enabled) {
  PF_API_BEGIN
  if (!B) fail(PF_EINVAL, "Batch: null handle");
  B->use_graph = enabled != 0;
  if (B->graph) {
    PF_HIP(hipStreamSynchronize(B->ctx->stream));  // not destroyed while a replay runs
    (void)hipGraphExecDestroy(B->graph);
    B->graph = nullptr;
  }
  B->graph_valid = false;
  PF_API_END
}

int pf_batch_destroy(pf_batch* B) {
  PF_API_BEGIN
  if (!B) return PF_OK;
  (void)hipStreamSynchronize(B->ctx->stream);
  if (B->graph) (void)hipGraphExecDestroy(B->graph);
  for (auto& g : B->groups) batch_group_free(g);
  dfree(B->dev);
  dfree(B->scan_ctl);
  delete B;
  PF_API_END
}

// ---------------------------------------------------------------------------
// Scene layout (Simulation.ts:957-1104, layout.ts:7-42)
// ---------------------------------------------------------------------------
static uint32_t fill_box(const double mn[3], const double mx[3], double sp, float* out, uint32_t cap) {
  uint32_t k = 0;
  if (!(mn[0] <= mx[0] && mn[1] <= mx[1] && mn[2] <= mx[2])) return 0;  // empty box
  for (double y = mn[1] + sp / 2; y < mx[1]; y += sp)
    for (double z = mn[2] + sp / 2; z < mx[2]; z += sp)
      for (double x = mn[0] + sp / 2; x < mx[0]; x += sp) {
        if (out && k < cap) {
          out[3 * (size_t)k] = (float)x;
          out[3 * (size_t)k + 1] = (float)y;
          out[3 * (size_t)k + 2] = (float)z;
        }
        ++k;
      }
  return k;
}

int pf_layout_fill_box(const double mn[3], const double mx[3], double spacing, float* out, uint32_t cap,
                       uint32_t* count) {
  PF_API_BEGIN
  if (!(spacing > 0)) fail(PF_EINVAL, "fillBox: spacing must be positive, got %g", spacing);
  const uint32_t k = fill_box(mn, mx, spacing, out, cap);
  if (count) *count = k;
  PF_API_END
}

double pf_layout_radius_for_budget(double volume, double area, double cloth_spacing, uint32_t budget) {
  auto count = [&](double r) { return volume / std::pow(2 * r, 3) + area / std::pow(cloth_spacing * r, 2); };
  double low = 1e-5, high = 10;
  for (int i = 0; i < 80; ++i) {
    const double mid = std::sqrt(low * high);
    if (count(mid) > budget) low = mid;
    else high = mid;
  }
  return high;
}

// Box3.intersect and expandByScalar(-r), then fillBox at spacing 2r.
static uint32_t layout_count(const double* container, const double* boxes, uint32_t nb, double r) {
  uint32_t total = 0;
  for (uint32_t b = 0; b < nb; ++b) {
    double mn[3], mx[3];
    for (int a = 0; a < 3; ++a) {
      mn[a] = boxes[6 * b + a];
      mx[a] = boxes[6 * b + 3 + a];
      if (container) {
        mn[a] = std::max(mn[a], container[a] + r);
        mx[a] = std::min(mx[a], container[3 + a] - r);
      }
    }
    total += fill_box(mn, mx, 2 * r, nullptr, 0);
  }
  return total;
}

int pf_layout_fit_budget(uint32_t budget, const double container[6], const double* boxes, uint32_t nb,
                         double* radius, uint32_t* count) {
  PF_API_BEGIN
  if (budget == 0) fail(PF_EINVAL, "Simulation: particles must be a positive integer, got %u", budget);
  if (!boxes || nb == 0) fail(PF_EINVAL, "Simulation: add a fluid, smoke, soft body, or cloth before stepping");
  double volume = 0;
  for (uint32_t b = 0; b < nb; ++b) {
    double sz[3];
    bool empty = false;
    for (int a = 0; a < 3; ++a) {
      double lo = boxes[6 * b + a], hi = boxes[6 * b + 3 + a];
      if (container) { lo = std::max(lo, container[a]); hi = std::min(hi, container[3 + a]); }
      sz[a] = hi - lo;
      if (hi < lo) empty = true;
    }
    if (!empty) volume += sz[0] * sz[1] * sz[2];
  }
  const double guess = pf_layout_radius_for_budget(volume, 0, 2.2, budget);
  double r = guess;
  bool have = false;
  double best_r = 0;
  uint32_t best_n = 0;
  for (int pass = 0; pass < 12; ++pass) {
    const uint32_t c = layout_count(container, boxes, nb, r);
    if (c <= budget && (!have || c > best_n)) { have = true; best_r = r; best_n = c; }
    if (c <= budget && c >= 0.97 * budget) break;
    r *= std::cbrt(std::max<double>(c, 1) / (0.985 * budget));
  }
  for (int pass = 0; !have && pass < 100; ++pass) {
    r *= 1.05;
    const uint32_t c = layout_count(container, boxes, nb, r);
    if (c <= budget) { have = true; best_r = r; best_n = c; }
  }
  if (!have) fail(PF_EINVAL, "Simulation: the scene can't fit in %u particles", budget);
  if (radius) *radius = best_r;
  if (count) *count = best_n;
  PF_API_END
}

}  // extern "C"
