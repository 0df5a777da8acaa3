// Serving loop: resident scenes, a FIFO request queue, admission by a device
// memory budget, and continuous batching over several streams.
//
// A worker thread owns every HIP call after creation.  Each pass it retires
// batches whose completion event has fired (their requests become done), then
// for every idle stream takes queued requests in arrival order while their
// outputs fit the budget, and launches them as one batched render.  A request
// that arrives while batches run therefore starts as soon as a stream frees
// up, overlapping the batches still on the device.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <mutex>
#include <new>
#include <thread>
#include <unordered_map>
#include <vector>

#include "csg.h"
#include "device.hpp"
#include "runtime.hpp"

using namespace csgray::rt;

namespace {

using Clock = std::chrono::steady_clock;

size_t Align256(size_t n) { return (n + 255) & ~static_cast<size_t>(255); }

// Output block of one request: rgba | depth | id | flags, each 256-B aligned.
size_t OutputBytes(int64_t pixels) {
  return Align256(16 * pixels) + Align256(4 * pixels) + Align256(4 * pixels) + Align256(pixels);
}

enum State { kQueued = 0, kRunning = 1, kDone = 2 };

struct Request {
  csg_request req{};
  size_t bytes = 0;
  void* mem = nullptr;
  csg_outputs out{};
  State state = kQueued;
  csg_request_stats stats{};
};

struct Slot {
  hipStream_t stream = nullptr;
  hipEvent_t done = nullptr;
  std::vector<int64_t> tickets;  // empty = idle
};

}  // namespace

struct csg_server {
  csg_server_config cfg{};
  int device = 0;
  Clock::time_point t0;
  std::vector<const csg_scene*> scenes;
  std::mutex mu;
  std::condition_variable wake;   // worker: new work or shutdown
  std::condition_variable doneCv; // clients: a request finished
  std::deque<int64_t> queue;
  std::unordered_map<int64_t, Request> reqs;
  int64_t nextTicket = 1;
  int64_t used = 0, cached = 0;
  std::multimap<size_t, void*> cache;  // released output blocks by size
  std::vector<Slot> slots;
  csg_server_stats stats{};
  bool stop = false;
  std::thread worker;
  csgray::LaunchConfig launch;
};

namespace {

double Since(const csg_server* s) { return std::chrono::duration<double>(Clock::now() - s->t0).count(); }

// Takes an output block of `bytes` from the cache or the device; the caller
// holds the lock and has checked the budget.
void* TakeBlock(csg_server* s, size_t bytes) {
  auto it = s->cache.find(bytes);
  if (it != s->cache.end()) {
    void* p = it->second;
    s->cache.erase(it);
    s->cached -= static_cast<int64_t>(bytes);
    return p;
  }
  // make room: cached blocks of other sizes count against the budget too
  while (!s->cache.empty() && s->used + s->cached + static_cast<int64_t>(bytes) > s->cfg.budget_bytes) {
    auto victim = s->cache.begin();
    s->cached -= static_cast<int64_t>(victim->first);
    (void)hipFree(victim->second);
    s->cache.erase(victim);
  }
  void* p = nullptr;
  if (hipMalloc(&p, bytes) != hipSuccess) return nullptr;
  return p;
}

void Retire(csg_server* s) {
  bool any = false;
  for (Slot& sl : s->slots) {
    if (sl.tickets.empty()) continue;
    if (hipEventQuery(sl.done) != hipSuccess) continue;
    const double t = Since(s);
    for (int64_t tk : sl.tickets) {
      auto it = s->reqs.find(tk);
      if (it == s->reqs.end()) continue;
      it->second.state = kDone;
      it->second.stats.done_s = t;
      s->stats.completed++;
      s->stats.rays_done += static_cast<int64_t>(it->second.req.width) * it->second.req.height;
    }
    sl.tickets.clear();
    any = true;
  }
  if (any) s->doneCv.notify_all();
}

// Admits queued requests for one idle slot and launches them; lock held.
void Dispatch(csg_server* s, Slot& sl) {
  std::vector<int64_t> take;
  while (!s->queue.empty() && static_cast<int32_t>(take.size()) < s->cfg.max_batch_jobs) {
    Request& r = s->reqs[s->queue.front()];
    if (s->used + static_cast<int64_t>(r.bytes) > s->cfg.budget_bytes) break;  // wait for releases
    void* mem = TakeBlock(s, r.bytes);
    if (!mem) break;
    s->used += static_cast<int64_t>(r.bytes);
    s->stats.peak_bytes_in_use = std::max(s->stats.peak_bytes_in_use, s->used);
    r.mem = mem;
    const int64_t px = static_cast<int64_t>(r.req.width) * r.req.height;
    char* b = static_cast<char*>(mem);
    r.out.rgba = reinterpret_cast<float*>(b);
    r.out.depth = reinterpret_cast<float*>(b + Align256(16 * px));
    r.out.prim_id = reinterpret_cast<int32_t*>(b + Align256(16 * px) + Align256(4 * px));
    r.out.flags = reinterpret_cast<uint8_t*>(b + Align256(16 * px) + 2 * Align256(4 * px));
    take.push_back(s->queue.front());
    s->queue.pop_front();
  }
  if (take.empty()) return;
  std::vector<csgray::Job> jobs(take.size());
  int64_t base = 0;
  int32_t depth = 1;
  csg_render_opts opts{};
  opts.mode = s->cfg.mode;
  const double t = Since(s);
  for (size_t i = 0; i < take.size(); i++) {
    Request& r = s->reqs[take[i]];
    const csg_scene* sc = s->scenes[r.req.scene_id];
    (void)BuildJob(sc, &r.req.camera, &r.req.light, &r.req.phong, r.req.width, r.req.height, &r.out, &opts,
                   s->device, base, &jobs[i]);
    base += jobs[i].tiles;
    depth = std::max(depth, sc->dev.opDepth);
    r.state = kRunning;
    r.stats.start_s = t;
    r.stats.batch_jobs = static_cast<int32_t>(take.size());
  }
  csgray::RenderParams p{};
  csg_status st = UploadJobs(s->device, sl.stream, jobs.data(), jobs.size(), &p.jobs);
  csgray::LaunchShape shape;
  if (st == CSG_OK &&
      csgray::ChooseShape(depth, 0, s->launch, &shape, static_cast<int32_t>(jobs.size())) != hipSuccess)
    st = CSG_E_HIP;
  uint32_t* queue = nullptr;
  if (st == CSG_OK) st = QueueFor(s->device, sl.stream, &queue);
  uint2* heavy = nullptr;
  if (st == CSG_OK) st = HeavyFor(s->device, sl.stream, &heavy);
  if (st == CSG_OK) {
    p.numJobs = static_cast<int32_t>(jobs.size());
    p.numTiles = static_cast<int32_t>(base);
    p.stackSlots = depth;
    p.queue = queue;
    csgray::FitGridToTiles(&shape, base);
    p.totalWaves = static_cast<uint32_t>(shape.grid) * static_cast<uint32_t>(shape.block / 64);
    int32_t maxOps = 1, maxLeaves = 1;
    for (int64_t tk : take) {
      const csgray::DevScene& d = s->scenes[s->reqs[tk].req.scene_id]->dev;
      maxOps = std::max(maxOps, d.nOps);
      maxLeaves = std::max(maxLeaves, d.opLeaves);
    }
    PrepareCoop(s->launch, maxOps, maxLeaves, heavy, shape, &p);
    if (csgray::LaunchRender(p, s->launch, shape, sl.stream) != hipSuccess) st = CSG_E_HIP;
  }
  if (st == CSG_OK && hipEventRecord(sl.done, sl.stream) != hipSuccess) st = CSG_E_HIP;
  // A failed launch still retires its requests (outputs undefined) so that
  // no client blocks forever; the error is visible as a short batch time.
  if (st != CSG_OK) (void)hipEventRecord(sl.done, sl.stream);
  for (int64_t tk : take) s->reqs[tk].stats.status = st;
  sl.tickets = std::move(take);
  s->stats.batches++;
}

void WorkerMain(csg_server* s) {
  (void)hipSetDevice(s->device);
  std::unique_lock<std::mutex> lock(s->mu);
  for (;;) {
    Retire(s);
    bool inflight = false;
    for (Slot& sl : s->slots) {
      if (sl.tickets.empty() && !s->queue.empty()) Dispatch(s, sl);
      inflight |= !sl.tickets.empty();
    }
    if (s->stop && !inflight && s->queue.empty()) break;
    if (!inflight && s->queue.empty()) {
      s->wake.wait(lock, [s] { return s->stop || !s->queue.empty(); });
    } else {
      // poll the in-flight batches' events; short sleeps keep completion
      // times within ~20 us without holding the lock
      lock.unlock();
      std::this_thread::sleep_for(std::chrono::microseconds(20));
      lock.lock();
    }
  }
}

}  // namespace

extern "C" {

csg_status csg_server_create(const csg_server_config* cfg, csg_server** out) {
  if (!out) return Fail(CSG_E_ARG, "csg_server_create: out is NULL");
  *out = nullptr;
  csg_server* s = new (std::nothrow) csg_server();
  if (!s) return Fail(CSG_E_OOM, "csg_server_create: host allocation failed");
  if (cfg) s->cfg = *cfg;
  if (s->cfg.budget_bytes <= 0) s->cfg.budget_bytes = int64_t{16} << 30;
  if (s->cfg.streams <= 0) s->cfg.streams = 3;
  if (s->cfg.max_batch_jobs <= 0) s->cfg.max_batch_jobs = 256;
  s->cfg.max_batch_jobs = std::min(s->cfg.max_batch_jobs, csgray::kMaxBatchJobs);
  csg_render_opts o{};
  o.mode = s->cfg.mode;
  o.scene_placement = 1;
  o.coop_budget = s->cfg.coop_budget;
  o.coop_when = s->cfg.coop_when;
  o.coop_grace = s->cfg.coop_grace;
  csg_status st = ParseOpts(&o, &s->launch);
  if (st == CSG_OK && hipGetDevice(&s->device) != hipSuccess) st = Fail(CSG_E_HIP, "hipGetDevice failed");
  s->slots.resize(static_cast<size_t>(s->cfg.streams));
  for (Slot& sl : s->slots) {
    if (st != CSG_OK) break;
    if (hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&sl.done, hipEventDisableTiming) != hipSuccess)
      st = Fail(CSG_E_HIP, "stream or event creation failed");
  }
  if (st != CSG_OK) {
    for (Slot& sl : s->slots) {
      if (sl.done) (void)hipEventDestroy(sl.done);
      if (sl.stream) (void)hipStreamDestroy(sl.stream);
    }
    delete s;
    return st;
  }
  s->t0 = Clock::now();
  s->worker = std::thread(WorkerMain, s);
  *out = s;
  return CSG_OK;
}

csg_status csg_server_add_scene(csg_server* s, const csg_scene* scene, int32_t* id) {
  if (!s || !scene || !id) return Fail(CSG_E_ARG, "csg_server_add_scene: NULL argument");
  if (!scene->dmem || scene->device != s->device)
    return Fail(CSG_E_ARG, "csg_server_add_scene: scene is not on the server's device");
  std::lock_guard<std::mutex> lock(s->mu);
  s->scenes.push_back(scene);
  *id = static_cast<int32_t>(s->scenes.size() - 1);
  return CSG_OK;
}

csg_status csg_server_submit(csg_server* s, const csg_request* rq, int64_t* ticket) {
  if (!s || !rq || !ticket) return Fail(CSG_E_ARG, "csg_server_submit: NULL argument");
  if (rq->width < 2 || rq->height < 2) return Fail(CSG_E_ARG, "csg_server_submit: width and height must be >= 2");
  const int64_t px = static_cast<int64_t>(rq->width) * rq->height;
  if (px > (int64_t{1} << 31)) return Fail(CSG_E_ARG, "csg_server_submit: image too large");
  const size_t bytes = OutputBytes(px);
  std::lock_guard<std::mutex> lock(s->mu);
  if (rq->scene_id < 0 || rq->scene_id >= static_cast<int32_t>(s->scenes.size()))
    return Fail(CSG_E_ARG, "csg_server_submit: unknown scene id");
  if (static_cast<int64_t>(bytes) > s->cfg.budget_bytes) {
    s->stats.rejected++;
    return Fail(CSG_E_BUDGET, "csg_server_submit: the request's outputs exceed the memory budget");
  }
  const int64_t t = s->nextTicket++;
  Request& r = s->reqs[t];
  r.req = *rq;
  r.bytes = bytes;
  r.stats.submit_s = Since(s);
  s->queue.push_back(t);
  s->stats.submitted++;
  *ticket = t;
  s->wake.notify_one();
  return CSG_OK;
}

csg_status csg_server_poll(csg_server* s, int64_t ticket, int32_t* done) {
  if (!s || !done) return Fail(CSG_E_ARG, "csg_server_poll: NULL argument");
  std::lock_guard<std::mutex> lock(s->mu);
  auto it = s->reqs.find(ticket);
  if (it == s->reqs.end()) return Fail(CSG_E_ARG, "csg_server_poll: unknown ticket");
  *done = it->second.state == kDone ? 1 : 0;
  return CSG_OK;
}

csg_status csg_server_wait(csg_server* s, int64_t ticket, csg_outputs* out, csg_request_stats* stats) {
  if (!s) return Fail(CSG_E_ARG, "csg_server_wait: NULL server");
  std::unique_lock<std::mutex> lock(s->mu);
  auto it = s->reqs.find(ticket);
  if (it == s->reqs.end()) return Fail(CSG_E_ARG, "csg_server_wait: unknown ticket");
  s->doneCv.wait(lock, [&] { return it->second.state == kDone; });
  if (out) *out = it->second.out;
  if (stats) *stats = it->second.stats;
  return CSG_OK;
}

csg_status csg_server_copy(csg_server* s, int64_t ticket, const csg_outputs* dst, void* stream) {
  if (!s || !dst) return Fail(CSG_E_ARG, "csg_server_copy: NULL argument");
  csg_outputs src{};
  int64_t px = 0;
  {
    std::lock_guard<std::mutex> lock(s->mu);
    auto it = s->reqs.find(ticket);
    if (it == s->reqs.end()) return Fail(CSG_E_ARG, "csg_server_copy: unknown ticket");
    if (it->second.state != kDone) return Fail(CSG_E_ARG, "csg_server_copy: request not done");
    src = it->second.out;
    px = static_cast<int64_t>(it->second.req.width) * it->second.req.height;
  }
  const hipStream_t hs = static_cast<hipStream_t>(stream);
  const void* from[4] = {src.rgba, src.depth, src.prim_id, src.flags};
  void* to[4] = {dst->rgba, dst->depth, dst->prim_id, dst->flags};
  const int64_t sz[4] = {16 * px, 4 * px, 4 * px, px};
  for (int k = 0; k < 4; k++) {
    if (!to[k]) continue;
    const hipError_t e = hipMemcpyAsync(to[k], from[k], static_cast<size_t>(sz[k]), hipMemcpyDeviceToDevice, hs);
    if (e != hipSuccess) return FailHip(e, "hipMemcpyAsync of request outputs");
  }
  return CSG_OK;
}

csg_status csg_server_release(csg_server* s, int64_t ticket) {
  if (!s) return Fail(CSG_E_ARG, "csg_server_release: NULL server");
  std::lock_guard<std::mutex> lock(s->mu);
  auto it = s->reqs.find(ticket);
  if (it == s->reqs.end()) return Fail(CSG_E_ARG, "csg_server_release: unknown ticket");
  if (it->second.state != kDone) return Fail(CSG_E_ARG, "csg_server_release: request not done");
  if (it->second.mem) {
    s->cache.insert({it->second.bytes, it->second.mem});
    s->cached += static_cast<int64_t>(it->second.bytes);
    s->used -= static_cast<int64_t>(it->second.bytes);
  }
  s->reqs.erase(it);
  s->wake.notify_one();  // memory freed: queued requests may now fit
  return CSG_OK;
}

csg_status csg_server_stats_get(csg_server* s, csg_server_stats* st) {
  if (!s || !st) return Fail(CSG_E_ARG, "csg_server_stats_get: NULL argument");
  std::lock_guard<std::mutex> lock(s->mu);
  *st = s->stats;
  st->bytes_in_use = s->used;
  return CSG_OK;
}

csg_status csg_server_destroy(csg_server* s) {
  if (!s) return CSG_OK;
  {
    std::lock_guard<std::mutex> lock(s->mu);
    s->stop = true;
    s->wake.notify_one();
  }
  if (s->worker.joinable()) s->worker.join();
  for (Slot& sl : s->slots) {
    (void)hipStreamSynchronize(sl.stream);
    (void)hipEventDestroy(sl.done);
    (void)hipStreamDestroy(sl.stream);
  }
  for (auto& kv : s->reqs)
    if (kv.second.mem) (void)hipFree(kv.second.mem);
  for (auto& kv : s->cache) (void)hipFree(kv.second);
  delete s;
  return CSG_OK;
}

}  // extern "C"
