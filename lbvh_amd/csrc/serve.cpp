// Request server (include/lbvh_mi355x.h, "serving"): continuous batching of
// (scene, ray batch, optional refit) requests on one GPU.
//
// One worker thread owns every device resource.  Requests are admitted in
// arrival order while their device footprint fits under the memory cap and
// an in-flight slot is free (idle resident scenes are evicted, least recently
// used first, when a new scene needs room).  Admitted requests start on one
// of several lanes (a context with its own stream): the upload of the
// request's data, and the build of a new scene or the refit of a resident one.
// Every request whose lane work has finished joins the next traversal launch:
// one kernel over all of their trees and ray buffers.  Work on one scene
// keeps arrival order: a request starts only after every earlier request on
// its scene has started, and a refit only after every earlier request on its
// scene has been traced, so a result does not depend on what else is in
// flight.  Host data passes through pinned staging buffers so that no copy
// blocks the worker on device progress.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <new>
#include <string>
#include <thread>
#include <vector>

#include "engine.hpp"
#include "lbvh_mi355x.h"
#include "runtime.hpp"

using namespace lbvh_eng;
using namespace lbvh_rt;

namespace {

using Clock = std::chrono::steady_clock;

// Size-class pools: device buffers (reused without hipFree, which would
// synchronize the device) and pinned host staging buffers.
struct Pool
{
  bool host = false;
  std::map<uint64_t, std::vector<void*>> free_by_size;
  uint64_t cached = 0; // bytes held in the free lists

  static uint64_t size_class(uint64_t bytes)
  {
    uint64_t c = 4096;
    while (c < bytes)
      c <<= 1;
    return c;
  }
  void* get(uint64_t bytes, uint64_t* cls)
  {
    *cls = size_class(bytes);
    auto it = free_by_size.find(*cls);
    if (it != free_by_size.end() && !it->second.empty()) {
      void* p = it->second.back();
      it->second.pop_back();
      cached -= *cls;
      return p;
    }
    void* p = nullptr;
    hipError_t e = host ? hipHostMalloc(&p, *cls, hipHostMallocDefault) : hipMalloc(&p, *cls);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      return nullptr;
    }
    return p;
  }
  void put(void* p, uint64_t cls)
  {
    if (!p)
      return;
    free_by_size[cls].push_back(p);
    cached += cls;
  }
  uint64_t release_all()
  {
    uint64_t freed = 0;
    for (auto& kv : free_by_size)
      for (void* p : kv.second) {
        if (host)
          (void)hipHostFree(p);
        else
          (void)hipFree(p);
        freed += kv.first;
      }
    free_by_size.clear();
    cached = 0;
    return freed;
  }
};

struct Buf
{
  void* p = nullptr;
  uint64_t cls = 0;
};

// Traversal launches go to these streams in turn, so that a launch can
// start while the previous ones' longest rays still run (measured on the
// 160-request stream: 2.48x the plain path at 4,000 req/s with four streams,
// 1.94x with two, 1.87x with one).
constexpr int kTraceStreams = 4;

struct Scene
{
  uint64_t id = 0;
  uint64_t n = 0;
  lbvh_bvh* bvh = nullptr;
  float* d_tri = nullptr;
  uint64_t bytes = 0;     // device bytes held (triangles + tree)
  bool owns_tri = true;   // false: tri memory borrowed from device requests
  hipEvent_t ev_tree = nullptr;  // after the last build / refit
  // after the last traversal that read it, one per trace stream (a refit
  // waits on all of them)
  hipEvent_t ev_trace[kTraceStreams] = {};
  bool has_tree_ev = false, has_trace_ev[kTraceStreams] = {};
  uint64_t last_use = 0;
  uint32_t refs = 0;      // requests (queued or active) naming it
  bool built = false;     // written under the server lock
  bool failed = false;    // its build failed: requests naming it fail
};

enum class St
{
  Queued,
  Admitted,
  Started,
  Tracing,
  Done
};

struct Req
{
  uint64_t ticket = 0;
  lbvh_request r{};
  uint64_t scene = 0;
  Scene* sc = nullptr; // valid while the request holds a reference
  bool creates = false;
  bool refit = false;
  std::atomic<St> st{ St::Queued };
  lbvh_status status = LBVH_OK;
  std::string error;
  uint64_t footprint = 0; // device bytes reserved while active
  Buf d_rays, d_out, h_in, h_out;
  hipEvent_t ev_ready = nullptr;
  int launch = -1; // traversal launch it joined
  uint32_t coalesced = 0;
  double t_submit = 0, t_admit = 0, t_ready = 0, t_done = 0;
};

struct Launch
{
  hipEvent_t done = nullptr;
  Buf d_jobs, h_jobs;
  std::vector<Req*> reqs;
};

} // namespace

struct lbvh_server
{
  int device = 0;
  lbvh_server_opts opts{};
  std::vector<lbvh_ctx*> lanes;
  uint32_t next_lane = 0;
  hipStream_t trace_streams[kTraceStreams] = {};
  uint32_t next_trace = 0;
  Clock::time_point t0;

  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  bool stop = false;
  std::deque<Req*> inbox;             // submitted, not admitted (arrival order)
  std::map<uint64_t, Req*> reqs;      // every request not yet collected
  std::map<uint64_t, Scene*> scenes;  // resident (and being built) scenes
  uint64_t next_ticket = 1, next_scene = 1;
  uint64_t tick = 0;

  // worker-owned
  std::vector<Req*> active; // admitted, arrival order
  std::vector<Launch*> launches;
  Pool dev_pool, host_pool;
  uint64_t mem_used = 0, mem_cap = 0, mem_peak = 0;
  uint64_t lane_ws = 0; // workspace bytes of the lanes
  lbvh_server_stats stats{};
  std::thread worker;
};

namespace {

double
us_since(const lbvh_server* s)
{
  return std::chrono::duration<double, std::micro>(Clock::now() - s->t0).count();
}

uint64_t
tree_bytes_estimate(uint64_t n, uint32_t leaf_order)
{
  // nodes 32, parent links 8, flags 4, other end 4, centroid bounds, leaf
  // primitives 4 (primitive-id leaves)
  return n * (32 + 8 + 4 + 4) + (leaf_order == LBVH_LEAF_PRIMITIVE_ID ? 4 * n : 0) + 4096;
}

uint64_t
workspace_estimate(uint64_t n, uint32_t code_bits)
{
  const uint64_t kw = code_bits == 30 ? 4 : 8;
  return n * (12 + 2 * kw + 2 * 4 + 8) + (1u << 20);
}

void
fail_req(lbvh_server* s, Req* q, lbvh_status st, const std::string& msg)
{
  q->status = st;
  q->error = msg;
  (void)s;
}

// Device bytes a request needs while active (new scenes: the scene itself,
// which stays resident afterwards and is then accounted to the scene).
bool
on_device(const Req* q)
{
  return (q->r.flags & LBVH_REQUEST_DEVICE) != 0;
}

uint64_t
footprint(const lbvh_server* s, const Req* q)
{
  const bool dev = on_device(q);
  uint64_t b = dev ? 0 : Pool::size_class(q->r.n_rays * 32) + Pool::size_class(q->r.n_rays * 16);
  if (q->creates)
    b += (dev ? 0 : q->r.n_tri * 48) + tree_bytes_estimate(q->r.n_tri, s->opts.leaf_order) +
         workspace_estimate(q->r.n_tri, s->opts.code_bits);
  return b;
}

void
release_req_buffers(lbvh_server* s, Req* q)
{
  s->dev_pool.put(q->d_rays.p, q->d_rays.cls);
  s->dev_pool.put(q->d_out.p, q->d_out.cls);
  s->host_pool.put(q->h_in.p, q->h_in.cls);
  s->host_pool.put(q->h_out.p, q->h_out.cls);
  q->d_rays = q->d_out = q->h_in = q->h_out = Buf{};
}

void
destroy_scene(Scene* sc)
{
  if (sc->bvh)
    lbvh_bvh_destroy(sc->bvh);
  if (sc->d_tri && sc->owns_tri)
    (void)hipFree(sc->d_tri);
  if (sc->ev_tree)
    (void)hipEventDestroy(sc->ev_tree);
  for (int k = 0; k < kTraceStreams; k++)
    if (sc->ev_trace[k])
      (void)hipEventDestroy(sc->ev_trace[k]);
  delete sc;
}

// Frees idle resident scenes, least recently used first, until `need`
// more bytes fit (caller holds the lock).
bool
make_room(lbvh_server* s, uint64_t need)
{
  if (s->mem_used + s->dev_pool.cached + need <= s->mem_cap)
    return true;
  // cached pool buffers go first
  if (s->dev_pool.cached) {
    s->dev_pool.release_all();
    if (s->mem_used + need <= s->mem_cap)
      return true;
  }
  std::vector<Scene*> idle;
  for (auto& kv : s->scenes) {
    Scene* sc = kv.second;
    if (sc->refs == 0 && sc->built)
      idle.push_back(sc);
  }
  std::sort(idle.begin(), idle.end(),
            [](const Scene* a, const Scene* b) { return a->last_use < b->last_use; });
  for (Scene* sc : idle) {
    if (s->mem_used + need <= s->mem_cap)
      break;
    for (int k = 0; k < kTraceStreams; k++)
      if (sc->has_trace_ev[k])
        (void)hipEventSynchronize(sc->ev_trace[k]);
    s->mem_used -= std::min(s->mem_used, sc->bytes);
    s->scenes.erase(sc->id);
    destroy_scene(sc);
    s->stats.evictions++;
  }
  return s->mem_used + need <= s->mem_cap;
}

// Moves requests from the inbox to the active list while they fit (caller
// holds the lock).  Arrival order is kept: the first request that does not
// fit waits, and so does everything behind it.
void
admit(lbvh_server* s)
{
  while (!s->inbox.empty() && s->active.size() < s->opts.max_inflight) {
    Req* q = s->inbox.front();
    if (!q->creates && q->sc->failed) {
      s->inbox.pop_front();
      fail_req(s, q, LBVH_E_INVALID, "the scene's build failed");
      q->st = St::Done;
      q->t_done = us_since(s);
      if (q->sc->refs && --q->sc->refs == 0) {
        s->scenes.erase(q->sc->id);
        destroy_scene(q->sc);
      }
      q->sc = nullptr;
      s->cv_done.notify_all();
      continue;
    }
    const uint64_t need = footprint(s, q);
    if (need > s->mem_cap) {
      s->inbox.pop_front();
      fail_req(s, q, LBVH_E_OOM, "request needs more device memory than the server's cap");
      q->st = St::Done;
      q->t_done = us_since(s);
      if (q->sc->refs && --q->sc->refs == 0 && (q->creates || q->sc->failed)) {
        if (q->creates)
          q->sc->failed = true;
        s->scenes.erase(q->sc->id);
        destroy_scene(q->sc);
      } else if (q->creates) {
        q->sc->failed = true;
      }
      q->sc = nullptr;
      s->cv_done.notify_all();
      continue;
    }
    if (!make_room(s, need)) {
      if (!s->active.empty())
        break; // wait for active requests to finish
      // nothing is running and every idle scene is gone: it can never fit
      s->inbox.pop_front();
      fail_req(s, q, LBVH_E_OOM, "request does not fit the memory cap beside the scenes in use");
      q->st = St::Done;
      q->t_done = us_since(s);
      if (q->creates)
        q->sc->failed = true;
      if (q->sc->refs && --q->sc->refs == 0 && q->sc->failed) {
        s->scenes.erase(q->sc->id);
        destroy_scene(q->sc);
      }
      q->sc = nullptr;
      s->cv_done.notify_all();
      continue;
    }
    s->inbox.pop_front();
    s->mem_used += need;
    s->mem_peak = std::max(s->mem_peak, s->mem_used);
    q->footprint = need;
    q->st = St::Admitted;
    q->t_admit = us_since(s);
    s->active.push_back(q);
  }
}

bool
earlier_on_scene_all(const lbvh_server* s, const Req* q, St at_least)
{
  for (const Req* o : s->active) {
    if (o == q)
      break;
    if (o->scene == q->scene && o->st < at_least)
      return false;
  }
  return true;
}

// Starts a request on a lane: staging, upload, build or refit, then its
// ready event.  Runs on the worker without the lock.
void
start(lbvh_server* s, Req* q, Scene* sc)
{
  lbvh_ctx* lane = s->lanes[s->next_lane++ % s->lanes.size()];
  hipStream_t st = lane->stream;
  auto bad = [&](lbvh_status code, const std::string& msg) {
    fail_req(s, q, code, msg);
  };
  const uint64_t nr = q->r.n_rays;
  const uint64_t ntri = (q->creates || q->refit) ? q->r.n_tri : 0;
  // pinned staging of the inputs (rays, then triangles), outputs
  q->h_in = Buf{};
  if (!q->creates && !sc->bvh)
    return bad(LBVH_E_INVALID, "the scene's build failed");
  const bool dev = on_device(q);
  const float* tri_src = q->r.tri48;
  if (!dev) {
    const uint64_t in_bytes = nr * 32 + ntri * 48;
    if (in_bytes) {
      q->h_in.p = s->host_pool.get(in_bytes, &q->h_in.cls);
      if (!q->h_in.p)
        return bad(LBVH_E_OOM, "pinned staging allocation failed");
      if (nr)
        std::memcpy(q->h_in.p, q->r.rays32, nr * 32);
      if (ntri)
        std::memcpy(static_cast<char*>(q->h_in.p) + nr * 32, q->r.tri48, ntri * 48);
    }
    q->h_out.p = s->host_pool.get(nr * 16 + 16, &q->h_out.cls);
    q->d_rays.p = s->dev_pool.get(nr * 32 + 32, &q->d_rays.cls);
    q->d_out.p = s->dev_pool.get(nr * 16 + 16, &q->d_out.cls);
    if (!q->h_out.p || !q->d_rays.p || !q->d_out.p)
      return bad(LBVH_E_OOM, "request buffer allocation failed");
    if (nr)
      if (hipMemcpyAsync(q->d_rays.p, q->h_in.p, nr * 32, hipMemcpyHostToDevice, st) !=
          hipSuccess)
        return bad(LBVH_E_HIP, "ray upload failed");
    tri_src = reinterpret_cast<const float*>(static_cast<char*>(q->h_in.p) + nr * 32);
  }
  lbvh_build_opts o;
  lbvh_build_opts_init(&o);
  o.code_bits = s->opts.code_bits;
  o.leaf_order = s->opts.leaf_order;
  if (q->creates) {
    if (dev) {
      sc->d_tri = const_cast<float*>(tri_src);
      sc->owns_tri = false;
    } else {
      if (hipMalloc(&sc->d_tri, ntri * 48) != hipSuccess) {
        (void)hipGetLastError();
        return bad(LBVH_E_OOM, "scene allocation failed");
      }
      if (hipMemcpyAsync(sc->d_tri, tri_src, ntri * 48, hipMemcpyHostToDevice, st) != hipSuccess)
        return bad(LBVH_E_HIP, "triangle upload failed");
    }
    const uint64_t ws0 = lbvh_ctx_workspace_bytes(lane);
    // a workspace that must grow waits for the device (the old buffers are
    // freed), so it grows to the next power of two at once
    uint64_t want = 2;
    while (want < ntri)
      want <<= 1;
    const int key_bits = o.code_bits == 30 ? 30 : 60;
    lbvh_status rs = reserve_build(lane, want, int(o.code_bits), untuned_sort(key_bits, want),
                                   o.leaf_order == LBVH_LEAF_PRIMITIVE_ID);
    if (rs != LBVH_OK)
      return bad(rs, lbvh_ctx_last_error(lane));
    lbvh_status bs = lbvh_build_triangles_device(lane, sc->d_tri, ntri, &o, &sc->bvh);
    if (bs != LBVH_OK)
      return bad(bs, lbvh_ctx_last_error(lane));
    const uint64_t ws1 = lbvh_ctx_workspace_bytes(lane);
    sc->bytes = (sc->owns_tri ? ntri * 48 : 0) + sc->bvh->bytes;
    std::lock_guard<std::mutex> g(s->mu);
    s->lane_ws += ws1 - ws0;
  } else {
    if (sc->has_tree_ev)
      (void)hipStreamWaitEvent(st, sc->ev_tree, 0);
    if (q->refit) {
      for (int k = 0; k < kTraceStreams; k++)
        if (sc->has_trace_ev[k])
          (void)hipStreamWaitEvent(st, sc->ev_trace[k], 0);
      if (!dev && !sc->owns_tri) {
        // host data for a scene whose triangles are borrowed device memory:
        // borrowed memory is only read, so the scene takes a buffer of its own
        float* own = nullptr;
        if (hipMalloc(&own, ntri * 48) != hipSuccess) {
          (void)hipGetLastError();
          return bad(LBVH_E_OOM, "scene allocation failed");
        }
        sc->d_tri = own;
        sc->owns_tri = true;
        std::lock_guard<std::mutex> g(s->mu);
        sc->bytes += ntri * 48;
        s->mem_used += ntri * 48;
        s->mem_peak = std::max(s->mem_peak, s->mem_used);
      }
      if (dev && !sc->owns_tri) {
        sc->d_tri = const_cast<float*>(tri_src); // borrow the new triangles
      } else if (hipMemcpyAsync(sc->d_tri, tri_src, ntri * 48,
                                dev ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice,
                                st) != hipSuccess) {
        return bad(LBVH_E_HIP, "triangle upload failed");
      }
      lbvh_status rs = lbvh_refit_triangles_device(lane, sc->bvh, sc->d_tri);
      if (rs != LBVH_OK)
        return bad(rs, lbvh_ctx_last_error(lane));
    }
  }
  if (hipEventRecord(q->ev_ready, st) != hipSuccess)
    return bad(LBVH_E_HIP, "event record failed");
  if (q->creates || q->refit) {
    (void)hipEventRecord(sc->ev_tree, st);
    sc->has_tree_ev = true;
  }
  if (q->creates) {
    std::lock_guard<std::mutex> g(s->mu);
    sc->built = true;
  }
}

void
launch_traversal(lbvh_server* s, std::vector<Req*>& ready)
{
  auto* L = new Launch;
  const uint32_t nj = uint32_t(ready.size());
  L->h_jobs.p = s->host_pool.get(sizeof(TraceJob) * nj, &L->h_jobs.cls);
  L->d_jobs.p = s->dev_pool.get(sizeof(TraceJob) * nj, &L->d_jobs.cls);
  (void)hipEventCreateWithFlags(&L->done, hipEventDisableTiming);
  auto* jobs = static_cast<TraceJob*>(L->h_jobs.p);
  uint64_t total = 0;
  for (uint32_t k = 0; k < nj; k++) {
    Req* q = ready[k];
    Scene* sc = q->sc;
    const uint64_t nr = q->r.n_rays;
    jobs[k].nodes = sc->bvh->d_nodes;
    jobs[k].tris = sc->d_tri;
    if (on_device(q)) {
      jobs[k].rays = q->r.rays32;
      jobs[k].prim = q->r.prim;
      jobs[k].t = q->r.t;
      jobs[k].uv = q->r.uv;
    } else {
      auto* out = static_cast<uint32_t*>(q->d_out.p);
      jobs[k].rays = static_cast<const float*>(q->d_rays.p);
      jobs[k].prim = out;
      jobs[k].t = reinterpret_cast<float*>(out + nr);
      jobs[k].uv = reinterpret_cast<float*>(out + 2 * nr);
    }
    jobs[k].ray_begin = total;
    jobs[k].nprims = uint32_t(sc->n);
    jobs[k].pad = 0;
    total += nr;
    q->coalesced = nj;
  }
  const int ts = int(s->next_trace++ % kTraceStreams);
  hipStream_t st = s->trace_streams[ts];
  bool ok = hipMemcpyAsync(L->d_jobs.p, L->h_jobs.p, sizeof(TraceJob) * nj, hipMemcpyHostToDevice,
                           st) == hipSuccess;
  if (ok && total)
    ok = launch_trace_jobs(int(s->opts.isect_mode), static_cast<const TraceJob*>(L->d_jobs.p), nj,
                           total, st) == hipSuccess;
  for (uint32_t k = 0; ok && k < nj; k++) {
    Req* q = ready[k];
    if (!on_device(q))
      ok = hipMemcpyAsync(q->h_out.p, q->d_out.p, q->r.n_rays * 16, hipMemcpyDeviceToHost, st) ==
           hipSuccess;
  }
  if (ok)
    ok = hipEventRecord(L->done, st) == hipSuccess;
  const int id = int(s->launches.size());
  s->launches.push_back(L);
  for (Req* q : ready) {
    q->launch = id;
    q->st = St::Tracing;
    if (!ok)
      fail_req(s, q, LBVH_E_HIP, "traversal launch failed");
    Scene* sc = q->sc;
    (void)hipEventRecord(sc->ev_trace[ts], st);
    sc->has_trace_ev[ts] = true;
  }
  L->reqs = ready;
  s->stats.launches++;
  s->stats.coalesced_requests += nj;
  s->stats.rays += total;
  if (nj > s->stats.max_coalesced)
    s->stats.max_coalesced = nj;
}

// Copies results out of staging and retires a request (worker, no lock held
// for the copy).
void
finish(lbvh_server* s, Req* q)
{
  const uint64_t nr = q->r.n_rays;
  if (q->status == LBVH_OK && nr && !on_device(q)) {
    const auto* o = static_cast<const uint32_t*>(q->h_out.p);
    std::memcpy(q->r.prim, o, nr * 4);
    std::memcpy(q->r.t, o + nr, nr * 4);
    if (q->r.uv)
      std::memcpy(q->r.uv, o + 2 * nr, nr * 8);
  }
  release_req_buffers(s, q);
  std::lock_guard<std::mutex> g(s->mu);
  Scene* sc = q->sc;
  if (sc) {
    if (sc->refs)
      sc->refs--;
    sc->last_use = ++s->tick;
    if (q->creates) {
      // the scene stays resident: its actual bytes replace the estimate that
      // the request's footprint reserved for it
      if (q->status == LBVH_OK)
        s->mem_used += sc->bytes;
      else
        sc->failed = true;
    }
    if (sc->failed && sc->refs == 0) {
      s->scenes.erase(sc->id);
      destroy_scene(sc);
    }
    q->sc = nullptr;
  }
  s->mem_used -= std::min(s->mem_used, q->footprint);
  s->mem_used += s->lane_ws; // lane workspace grows once and stays
  s->lane_ws = 0;
  q->footprint = 0;
  if (q->ev_ready)
    (void)hipEventDestroy(q->ev_ready);
  q->ev_ready = nullptr;
  q->st = St::Done;
  q->t_done = us_since(s);
  if (q->status == LBVH_OK)
    s->stats.completed++;
  else
    s->stats.failed++;
  s->active.erase(std::find(s->active.begin(), s->active.end(), q));
  s->cv_done.notify_all();
  s->cv_work.notify_all();
}

void
worker_main(lbvh_server* s)
{
  (void)hipSetDevice(s->device);
  for (;;) {
    std::vector<std::pair<Req*, Scene*>> to_start;
    {
      std::unique_lock<std::mutex> lk(s->mu);
      if (s->stop && s->inbox.empty() && s->active.empty())
        break;
      admit(s);
      for (Req* q : s->active) {
        if (q->st != St::Admitted)
          continue;
        if (!earlier_on_scene_all(s, q, St::Started))
          continue;
        if (q->refit && !earlier_on_scene_all(s, q, St::Tracing))
          continue;
        q->st = St::Started;
        to_start.emplace_back(q, q->sc);
      }
    }
    bool progress = !to_start.empty();
    for (auto& qs : to_start)
      start(s, qs.first, qs.second);

    // every started request whose lane work finished joins one launch, up to
    // kMaxRaysPerCall rays in all (the rest join the next one)
    std::vector<Req*> ready;
    std::vector<Req*> failed;
    uint64_t ready_rays = 0;
    for (Req* q : s->active) {
      if (q->st != St::Started)
        continue;
      if (q->status != LBVH_OK) {
        failed.push_back(q);
        continue;
      }
      if (ready_rays + q->r.n_rays > kMaxRaysPerCall)
        continue;
      if (hipEventQuery(q->ev_ready) == hipSuccess) {
        q->t_ready = us_since(s);
        ready_rays += q->r.n_rays;
        ready.push_back(q);
      }
    }
    for (Req* q : failed) {
      q->st = St::Tracing;
      finish(s, q);
      progress = true;
    }
    if (!ready.empty()) {
      launch_traversal(s, ready);
      progress = true;
    }
    // completed launches
    for (size_t i = 0; i < s->launches.size(); i++) {
      Launch* L = s->launches[i];
      if (!L || hipEventQuery(L->done) != hipSuccess)
        continue;
      for (Req* q : L->reqs)
        finish(s, q);
      s->dev_pool.put(L->d_jobs.p, L->d_jobs.cls);
      s->host_pool.put(L->h_jobs.p, L->h_jobs.cls);
      (void)hipEventDestroy(L->done);
      delete L;
      s->launches[i] = nullptr;
      progress = true;
    }
    while (!s->launches.empty() && s->launches.back() == nullptr)
      s->launches.pop_back();
    if (!progress) {
      std::unique_lock<std::mutex> lk(s->mu);
      s->cv_work.wait_for(lk, std::chrono::microseconds(50));
    }
  }
}

} // namespace

extern "C" {

void
lbvh_server_opts_init(lbvh_server_opts* o)
{
  if (!o)
    return;
  std::memset(o, 0, sizeof(*o));
  o->memory_cap_bytes = 288ull << 30;
  o->lanes = 4;
  o->max_inflight = 64;
  o->code_bits = 63;
  o->leaf_order = LBVH_LEAF_PRIMITIVE_ID;
  o->isect_mode = LBVH_ISECT_REFERENCE;
}

lbvh_status
lbvh_server_create(int device, const lbvh_server_opts* opts, lbvh_server** out)
{
  if (!out)
    return LBVH_E_INVALID;
  *out = nullptr;
  lbvh_server_opts o;
  lbvh_server_opts_init(&o);
  if (opts)
    o = *opts;
  if ((o.code_bits != 30 && o.code_bits != 63) || o.leaf_order > 1 || o.isect_mode > 2 ||
      o.lanes == 0 || o.lanes > 16 || o.max_inflight == 0)
    return LBVH_E_INVALID;
  auto* s = new (std::nothrow) lbvh_server;
  if (!s)
    return LBVH_E_OOM;
  s->device = device;
  s->opts = o;
  s->host_pool.host = true;
  s->t0 = Clock::now();
  for (uint32_t i = 0; i < o.lanes; i++) {
    lbvh_ctx* c = nullptr;
    lbvh_status st = lbvh_ctx_create(device, &c);
    if (st != LBVH_OK) {
      for (auto* l : s->lanes)
        lbvh_ctx_destroy(l);
      delete s;
      return st;
    }
    // a request's latency must not include tuning runs
    lbvh_ctx_set_autotune(c, 0);
    s->lanes.push_back(c);
    if (o.reserve_prims >= 2) {
      // the power of two a build of reserve_prims would grow the lane to
      uint64_t want = 2;
      while (want < o.reserve_prims)
        want <<= 1;
      const int key_bits = o.code_bits == 30 ? 30 : 60;
      st = reserve_build(c, want, int(o.code_bits), untuned_sort(key_bits, want),
                         o.leaf_order == LBVH_LEAF_PRIMITIVE_ID);
      if (st != LBVH_OK) {
        for (auto* l : s->lanes)
          lbvh_ctx_destroy(l);
        delete s;
        return st;
      }
      s->lane_ws += lbvh_ctx_workspace_bytes(c);
    }
  }
  for (int k = 0; k < kTraceStreams; k++) {
    if (hipStreamCreateWithFlags(&s->trace_streams[k], hipStreamNonBlocking) != hipSuccess) {
      for (int j = 0; j < k; j++)
        (void)hipStreamDestroy(s->trace_streams[j]);
      for (auto* l : s->lanes)
        lbvh_ctx_destroy(l);
      delete s;
      return LBVH_E_HIP;
    }
  }
  size_t free_b = 0, total_b = 0;
  (void)hipMemGetInfo(&free_b, &total_b);
  s->mem_cap = o.memory_cap_bytes;
  if (free_b && s->mem_cap > uint64_t(free_b) - (uint64_t(free_b) >> 4))
    s->mem_cap = uint64_t(free_b) - (uint64_t(free_b) >> 4); // keep 1/16 of free memory spare
  s->mem_used = s->lane_ws; // reserved lane workspaces
  s->mem_peak = s->mem_used;
  s->lane_ws = 0;
  s->worker = std::thread(worker_main, s);
  *out = s;
  return LBVH_OK;
}

void
lbvh_server_destroy(lbvh_server* s)
{
  if (!s)
    return;
  {
    std::lock_guard<std::mutex> g(s->mu);
    s->stop = true;
  }
  s->cv_work.notify_all();
  if (s->worker.joinable())
    s->worker.join();
  (void)hipSetDevice(s->device);
  for (int k = 0; k < kTraceStreams; k++)
    (void)hipStreamSynchronize(s->trace_streams[k]);
  for (auto& kv : s->scenes)
    destroy_scene(kv.second);
  for (auto& kv : s->reqs)
    delete kv.second;
  s->dev_pool.release_all();
  s->host_pool.release_all();
  for (auto* l : s->lanes)
    lbvh_ctx_destroy(l);
  for (int k = 0; k < kTraceStreams; k++)
    (void)hipStreamDestroy(s->trace_streams[k]);
  delete s;
}

lbvh_status
lbvh_server_submit(lbvh_server* s, const lbvh_request* r, uint64_t* ticket, uint64_t* scene_out)
{
  if (!s || !r || !ticket)
    return LBVH_E_INVALID;
  if (r->n_rays && (!r->rays32 || !r->prim || !r->t))
    return LBVH_E_INVALID;
  if (r->n_rays > kMaxRaysPerCall)
    return LBVH_E_INVALID;
  if (r->scene == 0 && (!r->tri48 || r->n_tri < 2 || r->n_tri >= (uint64_t(1) << 29)))
    return LBVH_E_INVALID;
  if ((r->flags & LBVH_REQUEST_DEVICE) &&
      (misaligned(r->tri48, 16) || misaligned(r->rays32, 16) ||
       trace_outputs_misaligned(r->prim, r->t, r->uv)))
    return LBVH_E_INVALID;
  auto* q = new (std::nothrow) Req;
  if (!q)
    return LBVH_E_OOM;
  q->r = *r;
  if ((hipSetDevice(s->device) != hipSuccess) ||
      hipEventCreateWithFlags(&q->ev_ready, hipEventDisableTiming) != hipSuccess) {
    delete q;
    return LBVH_E_HIP;
  }
  std::lock_guard<std::mutex> g(s->mu);
  if (s->stop) {
    (void)hipEventDestroy(q->ev_ready);
    delete q;
    return LBVH_E_BUSY;
  }
  if (r->scene == 0) {
    auto* sc = new (std::nothrow) Scene;
    bool made = sc && hipEventCreateWithFlags(&sc->ev_tree, hipEventDisableTiming) == hipSuccess;
    for (int k = 0; made && k < kTraceStreams; k++)
      made = hipEventCreateWithFlags(&sc->ev_trace[k], hipEventDisableTiming) == hipSuccess;
    if (!made) {
      if (sc)
        destroy_scene(sc);
      (void)hipEventDestroy(q->ev_ready);
      delete q;
      return LBVH_E_OOM;
    }
    sc->id = s->next_scene++;
    sc->refs = 1;
    sc->n = r->n_tri;
    s->scenes[sc->id] = sc;
    q->scene = sc->id;
    q->sc = sc;
    q->creates = true;
  } else {
    auto it = s->scenes.find(r->scene);
    if (it == s->scenes.end()) {
      (void)hipEventDestroy(q->ev_ready);
      delete q;
      return LBVH_E_INVALID;
    }
    if ((r->tri48 && r->n_tri != it->second->n) || it->second->failed) {
      (void)hipEventDestroy(q->ev_ready);
      delete q;
      return LBVH_E_INVALID;
    }
    it->second->refs++;
    q->scene = r->scene;
    q->sc = it->second;
    q->refit = r->tri48 != nullptr;
  }
  q->ticket = s->next_ticket++;
  q->t_submit = us_since(s);
  s->reqs[q->ticket] = q;
  s->inbox.push_back(q);
  *ticket = q->ticket;
  if (scene_out)
    *scene_out = q->scene;
  s->cv_work.notify_all();
  return LBVH_OK;
}

static void
fill_info(const Req* q, lbvh_request_info* info)
{
  if (!info)
    return;
  std::memset(info, 0, sizeof(*info));
  info->scene = q->scene;
  info->status = q->status;
  info->coalesced_with = q->coalesced;
  info->submit_us = q->t_submit;
  info->admit_us = q->t_admit;
  info->ready_us = q->t_ready;
  info->done_us = q->t_done;
}

lbvh_status
lbvh_server_poll(lbvh_server* s, uint64_t ticket, lbvh_request_info* info)
{
  if (!s)
    return LBVH_E_INVALID;
  std::lock_guard<std::mutex> g(s->mu);
  auto it = s->reqs.find(ticket);
  if (it == s->reqs.end())
    return LBVH_E_INVALID;
  Req* q = it->second;
  if (q->st != St::Done)
    return LBVH_E_BUSY;
  fill_info(q, info);
  const lbvh_status st = q->status;
  s->reqs.erase(it);
  delete q;
  return st;
}

lbvh_status
lbvh_server_wait(lbvh_server* s, uint64_t ticket, lbvh_request_info* info)
{
  if (!s)
    return LBVH_E_INVALID;
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->reqs.find(ticket);
  if (it == s->reqs.end())
    return LBVH_E_INVALID;
  Req* q = it->second;
  s->cv_done.wait(lk, [&] { return q->st == St::Done; });
  fill_info(q, info);
  const lbvh_status st = q->status;
  s->reqs.erase(ticket);
  delete q;
  return st;
}

lbvh_status
lbvh_server_release_scene(lbvh_server* s, uint64_t scene)
{
  if (!s)
    return LBVH_E_INVALID;
  std::lock_guard<std::mutex> g(s->mu);
  auto it = s->scenes.find(scene);
  if (it == s->scenes.end())
    return LBVH_E_INVALID;
  if (it->second->refs)
    return LBVH_E_BUSY;
  // the worker frees it the next time it needs room; make it first in line
  it->second->last_use = 0;
  return LBVH_OK;
}

lbvh_status
lbvh_server_get_stats(lbvh_server* s, lbvh_server_stats* out)
{
  if (!s || !out)
    return LBVH_E_INVALID;
  std::lock_guard<std::mutex> g(s->mu);
  *out = s->stats;
  out->memory_cap_bytes = s->mem_cap;
  out->memory_used_bytes = s->mem_used;
  out->memory_peak_bytes = s->mem_peak;
  out->resident_scenes = 0;
  for (auto& kv : s->scenes)
    if (kv.second->built)
      out->resident_scenes++;
  return LBVH_OK;
}

} // extern "C"
