// Host-side launchers of the device kernels (one per kernel family).
#pragma once
#include <tuple>

#include "pf_internal.hpp"

namespace pf {

// Launch shape of the tile kernels (those that take their tiles through
// pf_for_tiles). 0 launches one workgroup per tile; a positive cap (a multiple
// of the 8 XCDs) launches at most that many workgroups, each walking the tile
// list grid-stride: the persistent form. enqueue_step sets it from its plan
// for the launches it makes on the calling thread; other callers get 0.
extern thread_local uint32_t tile_grid_cap;
inline uint32_t tile_blocks(uint32_t ntiles) {
  return (tile_grid_cap != 0u && tile_grid_cap < ntiles) ? tile_grid_cap : ntiles;
}
struct TileGridScope {
  uint32_t saved;
  explicit TileGridScope(uint32_t cap) : saved(tile_grid_cap) { tile_grid_cap = cap; }
  ~TileGridScope() { tile_grid_cap = saved; }
};

// Launches a tile kernel over ntiles tiles with `lds` bytes of dynamic LDS:
// `one` (its PERSIST = false instantiation) with a workgroup per tile, or
// `many` (PERSIST = true) with tile_blocks(ntiles) workgroups when the launch
// shape caps the grid below the tile count. Launch errors surface through
// hipGetLastError like a triple-chevron launch's.
template <typename... P, typename... A>
void launch_tiles(void (*one)(P...), void (*many)(P...), uint32_t ntiles, size_t lds, hipStream_t st,
                  A&&... args) {
  static_assert(sizeof...(P) == sizeof...(A), "launch_tiles: argument count differs from the kernel's");
  if (ntiles == 0) return;
  const uint32_t blocks = tile_blocks(ntiles);
  std::tuple<P...> vals(static_cast<P>(args)...);
  std::apply(
      [&](auto&... v) {
        void* argv[] = {static_cast<void*>(&v)...};
        (void)hipLaunchKernel(blocks < ntiles ? reinterpret_cast<const void*>(many)
                                              : reinterpret_cast<const void*>(one),
                              dim3(blocks), dim3(PF_BLOCK), argv, lds, st);
      },
      vals);
}

// grid.hip: a scene substep's counting sort (predict + hash, scan, place,
// stable rank + state gather from parity into parity ^ 1)
void launch_hash(hipStream_t st, const SceneDev* scenes, const Tile* tiles, uint32_t ntiles, int parity,
                 int xs_prev, int reset_step_flags);
void launch_scan(hipStream_t st, const SceneDev* scenes, const uint2* stiles, uint32_t nstiles,
                 void* ctl);
void launch_place(hipStream_t st, const SceneDev* scenes, const Tile* tiles, uint32_t ntiles, int parity);
void launch_rank(hipStream_t st, const SceneDev* scenes, const Tile* tiles, uint32_t ntiles, int parity,
                 int xs_prev, int xs_out);
// The standalone grid's rebuild as an LSD radix sort of its x[0] positions:
// key (cell index), start[0..H], sorted keys, permutation and (spos non-null)
// sorted positions; tmp_key/tmp_src are the passes' other buffers. `scratch`
// holds grid_radix_scratch_words(n, H) words: the per-block digit histograms
// and the first sorted position of each offsets range.
constexpr uint32_t RS_ITEMS = 16;                  // points per thread
constexpr uint32_t RS_TILE = PF_BLOCK * RS_ITEMS;  // points per block (4096)
constexpr uint32_t RS_BINS = 256;                  // 8-bit digits
constexpr uint32_t RS_RANGE = 4096;                // buckets per offsets workgroup
constexpr size_t grid_radix_scratch_words(uint32_t n, uint32_t H) {
  return (size_t)RS_BINS * ((n + RS_TILE - 1) / RS_TILE) + RS_BINS + ((size_t)H + RS_RANGE - 1) / RS_RANGE + 1;
}
void launch_grid_radix(hipStream_t st, const SceneDev* dev, const SceneDev& host, uint32_t* skey,
                       uint32_t* perm, float4* spos, uint32_t* scratch);

// fluid.hip. `cap`: tile-union entries of dynamic LDS per workgroup
// (PF_BLOCK..PF_TU_MAX); the neighbour build keeps unions within it.
void launch_neighbors(hipStream_t st, const SceneDev* sc, const Tile* tl, uint32_t ntiles, int fi,
                      int p, int c, bool normal, uint32_t cap, int snap);
void launch_impulses(hipStream_t st, const SceneDev* sc, const Tile* tl, uint32_t ntiles, int fi,
                     int p, int c, bool surface_tension, bool adhesion);
// applies the surface-tension halves scattered into the accumulator
void launch_impulse_apply(hipStream_t st, const SceneDev* sc, const Tile* tl, uint32_t ntiles, int fi, int p,
                          int c);
void launch_lambda(hipStream_t st, const SceneDev* sc, const Tile* tl, uint32_t ntiles, int fi,
                   int p, int c, uint32_t cap);
void launch_delta(hipStream_t st, const SceneDev* sc, const Tile* tl, uint32_t ntiles, int fi,
                  int p, int c, bool walls, bool advect, int first_iteration, int substep, uint32_t cap);
void launch_walls(hipStream_t st, const SceneDev* sc, const Tile* tl, uint32_t ntiles, int p, int c,
                  bool advect, int first_iteration, int substep);
void launch_advect(hipStream_t st, const SceneDev* sc, const Tile* tl, uint32_t ntiles, int p, int c);
void launch_reaction_apply(hipStream_t st, const SceneDev* sc, const Tile* tl, uint32_t ntiles, int fi, int p,
                           int c);
void launch_boundary_volume(hipStream_t st, const SceneDev* sc, const Tile* tl, uint32_t ntiles, int fi, int p,
                            int c, int dynamic);
void launch_velocity(hipStream_t st, const SceneDev* sc, const Tile* tl, uint32_t ntiles, int fi,
                     int p, bool vort, bool visc, bool boundary, uint32_t cap);
void launch_vorticity(hipStream_t st, const SceneDev* sc, const Tile* tl, uint32_t ntiles, int fi,
                      int p, int c, bool vort, bool visc, bool friction, int substep,
                      int touch_valid, uint32_t cap);
// the ViscositySolver of fluid fi: seed, `sweeps` Jacobi sweeps, apply
void launch_viscosity_solve(hipStream_t st, const SceneDev* sc, const Tile* tl, uint32_t ntiles, int fi,
                            int p, uint32_t sweeps);
void launch_friction(hipStream_t st, const SceneDev* sc, const Tile* tl, uint32_t ntiles, int p,
                     int c, int substep);

// contact.hip: pair search (slot tiles), then the per-contact passes over
// contact tiles (scene, first record) in a grid-stride loop
void launch_contact_generate(hipStream_t st, const SceneDev* sc, const Tile* tl, uint32_t ntiles, int p, int c);
void launch_contact_stabilize(hipStream_t st, const SceneDev* sc, const Tile* tl, uint32_t ntiles,
                              const Tile* ctl, uint32_t nctiles, int p, int c);
void launch_contact_solve(hipStream_t st, const SceneDev* sc, const Tile* ctl, uint32_t nctiles, int p, int c);
void launch_contact_friction(hipStream_t st, const SceneDev* sc, const Tile* ctl, uint32_t nctiles, int p);

// softbody.hip
void launch_shape_match(hipStream_t st, const SceneDev* sc, const uint2* bodies, uint32_t nbodies,
                        int p, int c, int first_iteration, int pull);

// util kernels (runtime.hip)
void launch_gather_by_id(hipStream_t st, const SceneDev* sc, int p, int c, uint32_t n, int what,
                         void* out);
void launch_scatter_upload(hipStream_t st, const SceneDev* sc, int p, int c, uint32_t start,
                           uint32_t count, const float4* pos, const float4* vel, const float* inv_mass,
                           const uint32_t* group, uint32_t mask);

void launch_local_predict(hipStream_t st, const SceneDev* sc, const uint2* tiles, uint32_t ntiles);
void launch_local_solve(hipStream_t st, const SceneDev* sc, const uint2* tiles, uint32_t ntiles, int p, int c,
                        int first_iteration);
void launch_local_advect(hipStream_t st, const SceneDev* sc, const uint2* tiles, uint32_t ntiles);
}  // namespace pf
