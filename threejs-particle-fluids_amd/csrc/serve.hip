// Served form: a request loop over live scenes. Requests queue as they
// arrive; a worker thread launches every queued request whose scene is idle
// as one multi-scene launch (continuous batching) on one of several streams,
// keeps scenes resident in HBM under a byte budget and evicts idle scenes to
// host memory through their arenas. No counterpart in the reference.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <thread>

#include "runtime.hpp"

using namespace pf;

namespace {

using Clock = std::chrono::steady_clock;

enum : uint32_t { REQ_PENDING = 0, REQ_LAUNCHED = 1, REQ_DONE = 2, REQ_FAILED = 3 };

struct ServedScene {
  pf_loop_s* loop = nullptr;
  bool live = false;     // attached to the server
  bool busy = false;     // a launch holding the scene is in flight
  bool picked = false;   // taken by the launch being formed
  uint64_t last_use = 0;
  uint64_t bytes = 0;    // arena bytes while resident
  uint64_t host = 0;     // host bytes while evicted
  uint32_t particles = 0;
  uint32_t queued = 0;   // requests not yet launched
};

struct Request {
  uint32_t scene = 0;
  double dt = 0;
  uint32_t status = REQ_PENDING, launch_scenes = 0;
  double t_submit = 0, t_launch = 0, t_done = 0;
};

// One launch in flight: its own stream and scan ticket, so launches on
// different slots overlap on the device, plus the tables it reads.
struct Slot {
  Context ctx;
  hipEvent_t done = nullptr;
  bool busy = false;
  std::vector<uint64_t> tickets;
  std::vector<uint32_t> scenes;
  char* d_blob = nullptr;  // descriptors | tiles | scan tiles | body tiles | local tiles
  char* h_blob = nullptr;  // pinned staging of the same
  size_t cap = 0;
};

size_t align256(size_t b) { return (b + 255) / 256 * 256; }

}  // namespace

struct pf_server_s {
  Context* ctx = nullptr;
  pf_server_options opt{};
  uint32_t substeps = 0, iterations = 0;
  uint32_t grid_cap = 0;  // launch shape of the tile kernels (pf_server_set_launch_shape)
  int has_fluid = -1;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::thread worker;
  bool stop = false;
  std::vector<ServedScene> scenes;
  std::vector<Request> requests;  // by ticket
  std::deque<uint64_t> pending;   // submission order
  std::vector<std::unique_ptr<Slot>> slots;
  Clock::time_point t0 = Clock::now();
  uint64_t launch_counter = 0;
  pf_server_stats st{};
  int err_code = 0;
  std::string err;
  std::string note;  // the last refused eviction or restore (serving goes on)

  double now() const { return std::chrono::duration<double>(Clock::now() - t0).count(); }
  bool any_busy() const {
    for (auto& s : slots)
      if (s->busy) return true;
    return false;
  }
  uint64_t budget() const { return opt.hbm_budget_bytes ? opt.hbm_budget_bytes : ~0ull; }
  uint64_t host_budget() const { return opt.host_budget_bytes ? opt.host_budget_bytes : ~0ull; }

  // Moves an idle scene to host memory. Refused (false, the scene stays
  // resident and intact) when its copy would exceed the host budget or the
  // host memory or a copy fails; such failures never end serving.
  bool evict(uint32_t id) {
    ServedScene& sc = scenes[id];
    Arena* a = sc.loop->ps->arena.get();
    if (!a->resident) return true;
    const uint64_t need = a->used_bytes();
    if (st.host_bytes + need > host_budget()) {
      st.evict_failures += 1;
      note = "Server: host memory for evicted scenes is at its budget";
      return false;
    }
    try {
      a->evict();
    } catch (const PfError&) {
      st.evict_failures += 1;
      note = pf_last_error();
      return false;
    } catch (const HipError&) {
      st.evict_failures += 1;
      note = pf_last_error();
      return false;
    }
    sc.host = need;
    st.host_bytes += need;
    st.resident_bytes -= sc.bytes;
    st.resident_scenes -= 1;
    st.evictions += 1;
    return true;
  }
  // Brings a scene back; throws (the scene stays evicted) when no device
  // memory can be mapped.
  void restore(uint32_t id) {
    ServedScene& sc = scenes[id];
    Arena* a = sc.loop->ps->arena.get();
    if (a->resident) return;
    a->restore();
    st.host_bytes -= sc.host;
    sc.host = 0;
    st.resident_bytes += sc.bytes;
    st.resident_scenes += 1;
    st.peak_resident_bytes = std::max(st.peak_resident_bytes, st.resident_bytes);
    st.restores += 1;
  }
  // Evicts idle scenes, least recently stepped first, until `need` more
  // bytes fit the budget. Scenes in flight or picked stay, and a scene whose
  // eviction is refused is passed over. Returns whether the room was made;
  // *refused tells whether a refused eviction stood in the way (waiting
  // would not help then).
  bool make_room(uint64_t need, int64_t keep, bool* refused_any = nullptr) {
    std::vector<uint32_t> refused;
    bool ok = true;
    while (st.resident_bytes + need > budget()) {
      int64_t victim = -1;
      for (uint32_t i = 0; i < scenes.size(); ++i) {
        const ServedScene& s = scenes[i];
        if (!s.live || s.busy || s.picked || (int64_t)i == keep || !s.loop->ps->arena->resident) continue;
        if (std::find(refused.begin(), refused.end(), i) != refused.end()) continue;
        if (victim < 0 || s.last_use < scenes[victim].last_use) victim = i;
      }
      if (victim < 0) {
        ok = false;
        break;
      }
      if (!evict((uint32_t)victim)) refused.push_back((uint32_t)victim);
    }
    if (refused_any) *refused_any = !refused.empty();
    return ok;
  }

  bool poll() {
    bool any = false;
    for (auto& sp : slots) {
      Slot& s = *sp;
      if (!s.busy) continue;
      const hipError_t e = hipEventQuery(s.done);
      if (e == hipErrorNotReady) continue;
      PF_HIP(e);
      const double t = now();
      for (uint64_t tk : s.tickets) {
        Request& r = requests[tk];
        r.status = REQ_DONE;
        r.t_done = t;
        st.completed += 1;
        st.scene_steps += 1;
        st.particle_substeps += (uint64_t)scenes[r.scene].particles * substeps;
      }
      for (uint32_t id : s.scenes) scenes[id].busy = false;
      s.tickets.clear();
      s.scenes.clear();
      s.busy = false;
      any = true;
    }
    if (any) cv_done.notify_all();
    return any;
  }

  void fail_request(uint64_t tk) {
    Request& r = requests[tk];
    if (r.status == REQ_PENDING) scenes[r.scene].queued -= 1;
    r.status = REQ_FAILED;
    r.t_done = now();
    st.failed += 1;
  }

  // Forms and enqueues one launch on `s`. Returns whether it launched.
  bool launch(Slot& s) {
    const uint32_t max_scenes = opt.max_batch_scenes ? opt.max_batch_scenes : 256u;
    std::vector<uint64_t> pick;
    uint64_t parts = 0;
    for (uint64_t tk : pending) {
      if (pick.size() >= max_scenes) break;
      Request& r = requests[tk];
      ServedScene& sc = scenes[r.scene];
      if (sc.busy || sc.picked) continue;  // one step per scene per launch, in order
      if (opt.max_batch_particles && !pick.empty() && parts + sc.particles > opt.max_batch_particles) continue;
      sc.picked = true;
      parts += sc.particles;
      pick.push_back(tk);
    }
    // Residency, oldest request first: an accepted scene is protected, and
    // younger picks may be evicted to make room for an older one (they wait
    // for a later launch), so no request starves behind younger ones.
    for (uint64_t tk : pick) scenes[requests[tk].scene].picked = false;
    std::vector<uint64_t> run, failed;
    std::string why;
    for (uint64_t tk : pick) {
      const uint32_t id = requests[tk].scene;
      ServedScene& sc = scenes[id];
      if (!sc.loop->ps->arena->resident) {
        if (sc.bytes > budget()) {
          failed.push_back(tk);
          why = "Server: a scene needs more device memory than the HBM budget";
          continue;
        }
        bool refused = false;
        if (!make_room(sc.bytes, id, &refused)) {
          if (refused) {
            // idle scenes cannot go to host memory: this request fails alone
            failed.push_back(tk);
            why = note;
          }
          continue;  // otherwise it waits for a launch in flight to free memory
        }
        try {
          restore(id);
        } catch (const PfError&) {
          failed.push_back(tk);  // this request fails alone; the scene stays in host memory
          why = note = pf_last_error();
          continue;
        } catch (const HipError&) {
          failed.push_back(tk);
          why = note = pf_last_error();
          continue;
        }
      }
      sc.picked = true;
      run.push_back(tk);
    }
    for (uint64_t tk : failed) {
      fail_request(tk);
      pending.erase(std::find(pending.begin(), pending.end(), tk));
    }
    if (!failed.empty()) {
      err = why;
      cv_done.notify_all();
    }
    if (run.empty()) return !failed.empty();

    PlanShape sh;
    sh.substeps = substeps;
    sh.iterations = iterations;
    sh.grid_cap = grid_cap;
    std::vector<SceneDev> descs;
    std::vector<Tile> tl, ctl;
    std::vector<uint2> stl, btl, ltl;
    descs.reserve(run.size());
    for (uint32_t i = 0; i < run.size(); ++i) {
      const Request& r = requests[run[i]];
      pf_loop_s* L = scenes[r.scene].loop;
      loop_prepare_step(L, r.dt);
      plan_merge(sh, loop_plan_shape(L));
      loop_append_tiles(L, i, tl, stl, btl, ltl, ctl);
      descs.push_back(L->ps->desc(L->ps->variant));
    }
    const size_t o_tl = align256(descs.size() * sizeof(SceneDev));
    const size_t o_st = o_tl + align256(tl.size() * sizeof(Tile));
    const size_t o_bt = o_st + align256(stl.size() * sizeof(uint2));
    const size_t o_lt = o_bt + align256(btl.size() * sizeof(uint2));
    const size_t o_ct = o_lt + align256(std::max<size_t>(ltl.size(), 1) * sizeof(uint2));
    const size_t total = o_ct + align256(std::max<size_t>(ctl.size(), 1) * sizeof(Tile));
    if (total > s.cap) {
      if (s.d_blob) PF_HIP(hipFree(s.d_blob));
      if (s.h_blob) PF_HIP(hipHostFree(s.h_blob));
      s.d_blob = nullptr;
      s.h_blob = nullptr;
      s.cap = 0;
      const size_t cap = std::max<size_t>(total + total / 2, 1u << 20);
      PF_HIP(hipMalloc(reinterpret_cast<void**>(&s.d_blob), cap));
      PF_HIP(hipHostMalloc(reinterpret_cast<void**>(&s.h_blob), cap, hipHostMallocDefault));
      s.cap = cap;
    }
    std::memcpy(s.h_blob, descs.data(), descs.size() * sizeof(SceneDev));
    std::memcpy(s.h_blob + o_tl, tl.data(), tl.size() * sizeof(Tile));
    std::memcpy(s.h_blob + o_st, stl.data(), stl.size() * sizeof(uint2));
    std::memcpy(s.h_blob + o_bt, btl.data(), btl.size() * sizeof(uint2));
    if (!ltl.empty()) std::memcpy(s.h_blob + o_lt, ltl.data(), ltl.size() * sizeof(uint2));
    if (!ctl.empty()) std::memcpy(s.h_blob + o_ct, ctl.data(), ctl.size() * sizeof(Tile));
    PF_HIP(hipMemcpyAsync(s.d_blob, s.h_blob, total, hipMemcpyHostToDevice, s.ctx.stream));
    enqueue_step(&s.ctx, reinterpret_cast<const SceneDev*>(s.d_blob), sh,
                 reinterpret_cast<const Tile*>(s.d_blob + o_tl), (uint32_t)tl.size(),
                 reinterpret_cast<const uint2*>(s.d_blob + o_st), (uint32_t)stl.size(),
                 reinterpret_cast<const uint2*>(s.d_blob + o_bt), (uint32_t)btl.size(),
                 reinterpret_cast<const uint2*>(s.d_blob + o_lt), (uint32_t)ltl.size(),
                 reinterpret_cast<const Tile*>(s.d_blob + o_ct), (uint32_t)ctl.size(), s.ctx.scan_ctl);
    PF_HIP(hipGetLastError());
    PF_HIP(hipEventRecord(s.done, s.ctx.stream));

    const int flips = step_variant_flips(sh);
    const double t = now();
    launch_counter += 1;
    s.busy = true;
    for (uint64_t tk : run) {
      Request& r = requests[tk];
      ServedScene& sc = scenes[r.scene];
      loop_finish_step(sc.loop, flips);
      sc.busy = true;
      sc.picked = false;
      sc.queued -= 1;
      sc.last_use = launch_counter;
      r.status = REQ_LAUNCHED;
      r.t_launch = t;
      r.launch_scenes = (uint32_t)run.size();
      s.tickets.push_back(tk);
      s.scenes.push_back(r.scene);
    }
    for (uint64_t tk : pick) scenes[requests[tk].scene].picked = false;
    std::deque<uint64_t> rest;
    for (uint64_t tk : pending)
      if (requests[tk].status == REQ_PENDING) rest.push_back(tk);
    pending.swap(rest);
    st.launches += 1;
    st.max_launch_scenes = std::max<uint32_t>(st.max_launch_scenes, (uint32_t)run.size());
    return true;
  }

  // A failure of the worker ends serving: every queued or launched request
  // fails with its message and later calls report it.
  void fail_all(int code, const std::string& msg) {
    err_code = code;
    err = msg;
    for (uint64_t tk : pending) fail_request(tk);
    pending.clear();
    for (auto& sp : slots) {
      Slot& s = *sp;
      (void)hipStreamSynchronize(s.ctx.stream);
      for (uint64_t tk : s.tickets) fail_request(tk);
      for (uint32_t id : s.scenes) scenes[id].busy = false;
      s.tickets.clear();
      s.scenes.clear();
      s.busy = false;
    }
    for (auto& sc : scenes) sc.picked = false;
    cv_done.notify_all();
  }

  void run_worker() {
    std::unique_lock<std::mutex> lk(mu);
    while (true) {
      try {
        const bool progressed = poll();
        if (stop && pending.empty() && !any_busy()) break;
        bool launched = false;
        if (!pending.empty() && err_code == 0) {
          for (auto& sp : slots)
            if (!sp->busy) {
              launched = launch(*sp);
              break;
            }
        }
        if (progressed || launched) continue;
        if (any_busy())
          cv_work.wait_for(lk, std::chrono::microseconds(20));
        else
          cv_work.wait_for(lk, std::chrono::milliseconds(5));
      } catch (const PfError& e) {
        fail_all(e.code, pf_last_error());
      } catch (const HipError& e) {
        fail_all(e.code, pf_last_error());
      } catch (const std::exception& e) {
        fail_all(PF_EINVAL, e.what());
      }
    }
  }

  ServedScene& scene_at(uint32_t id, const char* what) {
    if (id >= scenes.size() || !scenes[id].live) fail(PF_EINVAL, "%s: no scene %u on this server", what, id);
    return scenes[id];
  }
  void check_error() {
    if (err_code) fail(err_code, "%s", err.c_str());
  }
};

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

extern "C" {

int pf_server_create(pf_context* ctx, const pf_server_options* opts, pf_server** out) {
  PF_API_BEGIN
  if (!ctx || !out) fail(PF_EINVAL, "Server: null argument");
  std::unique_ptr<pf_server_s> S(new pf_server_s());
  S->ctx = ctx;
  if (opts) S->opt = *opts;
  const uint32_t nslots = S->opt.slots ? S->opt.slots : 2u;
  if (nslots > 8) fail(PF_EINVAL, "Server: slots must be at most 8, got %u", nslots);
  S->st.hbm_budget_bytes = S->opt.hbm_budget_bytes;
  PF_HIP(hipSetDevice(ctx->device));
  for (uint32_t k = 0; k < nslots; ++k) {
    std::unique_ptr<Slot> s(new Slot());
    s->ctx.device = ctx->device;
    PF_HIP(hipStreamCreateWithFlags(&s->ctx.stream, hipStreamNonBlocking));
    PF_HIP(hipEventCreateWithFlags(&s->done, hipEventDisableTiming));
    s->ctx.scan_ctl = dalloc_temp<uint32_t>(4);
    S->slots.push_back(std::move(s));
  }
  pf_server_s* raw = S.get();
  S->worker = std::thread([raw] {
    (void)hipSetDevice(raw->ctx->device);
    raw->run_worker();
  });
  *out = S.release();
  PF_API_END
}

int pf_server_add_scene(pf_server* S, pf_loop* L, uint32_t* scene_id) {
  PF_API_BEGIN
  if (!S || !L || !scene_id) fail(PF_EINVAL, "Server: null argument");
  if (L->disposed || L->ps->disposed) fail(PF_ESTATE, "Server: the loop has been disposed");
  if (L->ctx != S->ctx) fail(PF_EINVAL, "Server: the loop must use the server's context");
  if (!L->ps->arena) fail(PF_EINVAL, "Server: the scene's particles must be created evictable");
  if (L->ps->server) fail(PF_EINVAL, "Server: the scene is already served");
  const int has_fluid = loop_has_grid(L) ? 1 : 0;
  PF_HIP(hipStreamSynchronize(L->ctx->stream));  // the owner's last plain step
  std::unique_lock<std::mutex> lk(S->mu);
  S->check_error();
  if (S->has_fluid < 0) {
    S->substeps = L->substeps;
    S->iterations = L->iterations;
    S->has_fluid = has_fluid;
  } else {
    if (L->substeps != S->substeps || L->iterations != S->iterations)
      fail(PF_EINVAL, "Server: every scene must share substeps and iterations (%u x %u)", S->substeps,
           S->iterations);
    if (has_fluid != S->has_fluid)
      fail(PF_EINVAL, "Server: mixing scenes with and without a neighbour grid (fluids or contacts) is not supported");
  }
  ServedScene sc;
  sc.loop = L;
  sc.live = true;
  sc.bytes = L->ps->arena->bytes();
  sc.particles = L->ps->n;
  sc.last_use = S->launch_counter;
  const uint32_t id = (uint32_t)S->scenes.size();
  S->scenes.push_back(sc);
  L->ps->server = S;
  S->st.scenes += 1;
  if (L->ps->arena->resident) {
    S->st.resident_bytes += sc.bytes;
    S->st.resident_scenes += 1;
  }
  // Admission: the new scene arrives resident (its owner built it); others
  // make room, else it waits in host memory. The cap holds from here on: a
  // scene that can neither fit nor go to host memory is not admitted.
  if (!S->make_room(0, id) && !S->evict(id)) {
    const std::string why = S->note;
    S->scenes.pop_back();
    L->ps->server = nullptr;
    S->st.scenes -= 1;
    S->st.resident_bytes -= sc.bytes;
    S->st.resident_scenes -= 1;
    fail(PF_ENOMEM, "Server: no room for the scene within the HBM budget, and it cannot be evicted (%s)", why.c_str());
  }
  S->st.peak_resident_bytes = std::max(S->st.peak_resident_bytes, S->st.resident_bytes);
  *scene_id = id;
  PF_API_END
}

static void wait_scene_idle(pf_server_s* S, std::unique_lock<std::mutex>& lk, uint32_t id) {
  S->cv_done.wait(lk, [&] { return S->err_code != 0 || (S->scenes[id].queued == 0 && !S->scenes[id].busy); });
}

int pf_server_remove_scene(pf_server* S, uint32_t id) {
  PF_API_BEGIN
  if (!S) fail(PF_EINVAL, "Server: null handle");
  std::unique_lock<std::mutex> lk(S->mu);
  ServedScene& sc0 = S->scene_at(id, "Server.removeScene");
  (void)sc0;
  wait_scene_idle(S, lk, id);
  ServedScene& sc = S->scenes[id];
  if (sc.busy) S->check_error();
  S->restore(id);
  sc.live = false;
  sc.loop->ps->server = nullptr;
  sc.loop->ps->dirty = true;
  S->st.scenes -= 1;
  S->st.resident_bytes -= sc.bytes;
  S->st.resident_scenes -= 1;
  S->check_error();
  PF_API_END
}

int pf_server_submit(pf_server* S, uint32_t id, double dt, uint64_t* ticket) {
  const double t = S ? S->now() : 0.0;
  PF_API_BEGIN
  if (!S) fail(PF_EINVAL, "Server: null handle");
  if (!(dt > 0) || !std::isfinite(dt)) fail(PF_EINVAL, "Server: step dt must be a positive finite number, got %g", dt);
  {
    std::lock_guard<std::mutex> lk(S->mu);
    S->check_error();
    if (S->stop) fail(PF_ESTATE, "Server: the server is shutting down");
    ServedScene& sc = S->scene_at(id, "Server.submit");
    Request r;
    r.scene = id;
    r.dt = dt;
    r.t_submit = t;
    const uint64_t tk = S->requests.size();
    S->requests.push_back(r);
    S->pending.push_back(tk);
    sc.queued += 1;
    S->st.submitted += 1;
    if (ticket) *ticket = tk;
  }
  S->cv_work.notify_one();
  PF_API_END
}

int pf_server_wait(pf_server* S, uint64_t ticket, pf_request_info* info) {
  PF_API_BEGIN
  if (!S) fail(PF_EINVAL, "Server: null handle");
  std::unique_lock<std::mutex> lk(S->mu);
  if (ticket >= S->requests.size()) fail(PF_EINVAL, "Server: unknown request %llu", (unsigned long long)ticket);
  S->cv_done.wait(lk, [&] { return S->requests[ticket].status >= REQ_DONE; });
  const Request& r = S->requests[ticket];
  if (info) *info = pf_request_info{r.scene, r.status, r.launch_scenes, 0, r.t_submit, r.t_launch, r.t_done};
  if (r.status == REQ_FAILED) fail(S->err_code ? S->err_code : PF_ENOMEM, "%s", S->err.empty() ? "Server: request failed" : S->err.c_str());
  PF_API_END
}

int pf_server_drain(pf_server* S) {
  PF_API_BEGIN
  if (!S) fail(PF_EINVAL, "Server: null handle");
  std::unique_lock<std::mutex> lk(S->mu);
  S->cv_done.wait(lk, [&] { return S->st.completed + S->st.failed == S->st.submitted; });
  S->check_error();
  PF_API_END
}

int pf_server_request_info(pf_server* S, uint64_t first, uint32_t count, pf_request_info* out) {
  PF_API_BEGIN
  if (!S || (count && !out)) fail(PF_EINVAL, "Server: null argument");
  std::lock_guard<std::mutex> lk(S->mu);
  if (first + count > S->requests.size()) fail(PF_EINVAL, "Server: request range out of bounds");
  for (uint32_t k = 0; k < count; ++k) {
    const Request& r = S->requests[first + k];
    out[k] = pf_request_info{r.scene, r.status, r.launch_scenes, 0, r.t_submit, r.t_launch, r.t_done};
  }
  PF_API_END
}

int pf_server_stats_get(pf_server* S, pf_server_stats* out) {
  PF_API_BEGIN
  if (!S || !out) fail(PF_EINVAL, "Server: null argument");
  std::lock_guard<std::mutex> lk(S->mu);
  *out = S->st;
  PF_API_END
}

double pf_server_now(pf_server* S) { return S ? S->now() : 0.0; }

int pf_server_set_launch_shape(pf_server* S, uint32_t workgroups) {
  PF_API_BEGIN
  if (!S) fail(PF_EINVAL, "Server: null handle");
  check_launch_shape(workgroups, "Server");
  std::lock_guard<std::mutex> lk(S->mu);
  S->grid_cap = workgroups;  // launches formed from now on
  PF_API_END
}

int pf_server_evict_scene(pf_server* S, uint32_t id) {
  PF_API_BEGIN
  if (!S) fail(PF_EINVAL, "Server: null handle");
  std::unique_lock<std::mutex> lk(S->mu);
  S->scene_at(id, "Server.evictScene");
  S->cv_done.wait(lk, [&] { return S->err_code != 0 || !S->scenes[id].busy; });
  S->check_error();
  if (!S->evict(id)) fail(PF_ENOMEM, "%s", S->note.c_str());
  PF_API_END
}

int pf_server_destroy(pf_server* S) {
  PF_API_BEGIN
  if (!S) return PF_OK;
  {
    std::lock_guard<std::mutex> lk(S->mu);
    S->stop = true;
  }
  S->cv_work.notify_all();
  if (S->worker.joinable()) S->worker.join();
  int code = PF_OK;
  std::string msg;
  for (uint32_t i = 0; i < S->scenes.size(); ++i) {
    ServedScene& sc = S->scenes[i];
    if (!sc.live) continue;
    try {
      S->restore(i);
    } catch (...) {
      if (code == PF_OK) {
        code = PF_EDEVICE;
        msg = pf_last_error();
      }
    }
    sc.live = false;
    sc.loop->ps->server = nullptr;
    sc.loop->ps->dirty = true;
  }
  for (auto& s : S->slots) {
    (void)hipStreamSynchronize(s->ctx.stream);
    (void)hipEventDestroy(s->done);
    dfree(s->ctx.scan_ctl);
    if (s->d_blob) (void)hipFree(s->d_blob);
    if (s->h_blob) (void)hipHostFree(s->h_blob);
    (void)hipStreamDestroy(s->ctx.stream);
  }
  delete S;
  if (code != PF_OK) fail(code, "%s", msg.c_str());
  PF_API_END
}

}  // extern "C"
