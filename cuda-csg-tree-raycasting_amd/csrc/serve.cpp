// Serving loop: resident scenes, a FIFO request queue, admission by a device
// memory budget, and continuous batching over several streams.
//
// A worker thread owns every launch after creation.  Each pass it retires
// batches whose completion event has fired (their requests become done), then
// for every idle stream takes queued requests in arrival order while their
// outputs fit the budget, and launches them as one batched render.  A request
// that arrives while batches run therefore starts as soon as a stream frees
// up, overlapping the batches still on the device.
//
// The budget covers everything the server holds on the device: each stream's
// workspace (tile queue, heavy-ray list with its resume area, job table),
// allocated when the server is created, and the requests' output blocks.
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

// An output block, and the event recorded after the last csg_server_copy out
// of it (null if none): the block is not written again before that copy.
struct Block {
  void* mem = nullptr;
  hipEvent_t copied = nullptr;
};

struct Request {
  csg_request req{};
  size_t bytes = 0;
  Block block;
  csg_outputs out{};
  State state = kQueued;
  csg_request_stats stats{};
};

struct Slot {
  hipStream_t stream = nullptr;
  hipEvent_t done = nullptr;
  std::vector<int64_t> tickets;      // empty = idle
  std::vector<hipEvent_t> waited;    // copy events this slot's batch waits on (recycled when it retires)
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
  int64_t used = 0, cached = 0;   // output bytes held by requests / in the cache
  int64_t workspace = 0;          // the streams' workspaces
  int64_t outputBudget = 0;       // budget_bytes - workspace
  std::multimap<size_t, Block> cache;  // released output blocks by size
  std::vector<hipEvent_t> eventPool;
  std::vector<Slot> slots;
  csg_server_stats stats{};
  bool stop = false;
  std::thread worker;
  csgray::LaunchConfig launch;
};

namespace {

double Since(const csg_server* s) { return std::chrono::duration<double>(Clock::now() - s->t0).count(); }

void ReturnEvent(csg_server* s, hipEvent_t e) {
  if (e) s->eventPool.push_back(e);
}

// Takes an output block of `bytes` from the cache or the device; the caller
// holds the lock and has checked the budget.
Block TakeBlock(csg_server* s, size_t bytes) {
  auto it = s->cache.find(bytes);
  if (it != s->cache.end()) {
    const Block b = it->second;
    s->cache.erase(it);
    s->cached -= static_cast<int64_t>(bytes);
    return b;
  }
  // make room: cached blocks of other sizes count against the budget too
  while (!s->cache.empty() && s->used + s->cached + static_cast<int64_t>(bytes) > s->outputBudget) {
    auto victim = s->cache.begin();
    s->cached -= static_cast<int64_t>(victim->first);
    if (victim->second.copied) (void)hipEventSynchronize(victim->second.copied);
    ReturnEvent(s, victim->second.copied);
    (void)hipFree(victim->second.mem);
    s->cache.erase(victim);
  }
  Block b;
  if (hipMalloc(&b.mem, bytes) != hipSuccess) b.mem = nullptr;
  return b;
}

// Marks a request done with `status`; the lock is held.
void Finish(csg_server* s, Request& r, csg_status status, double t) {
  r.state = kDone;
  r.stats.done_s = t;
  if (r.stats.status == CSG_OK) r.stats.status = status;
  s->stats.completed++;
  if (r.stats.status == CSG_OK) s->stats.rays_done += static_cast<int64_t>(r.req.width) * r.req.height;
}

void Retire(csg_server* s) {
  bool any = false;
  for (Slot& sl : s->slots) {
    if (sl.tickets.empty()) continue;
    const hipError_t q = hipEventQuery(sl.done);
    if (q == hipErrorNotReady) continue;
    // anything but "not ready" ends the batch; an error fails its requests
    const csg_status st = q == hipSuccess ? CSG_OK : CSG_E_HIP;
    const double t = Since(s);
    for (int64_t tk : sl.tickets) {
      auto it = s->reqs.find(tk);
      if (it != s->reqs.end()) Finish(s, it->second, st, t);
    }
    sl.tickets.clear();
    for (hipEvent_t e : sl.waited) ReturnEvent(s, e);
    sl.waited.clear();
    any = true;
  }
  if (any) s->doneCv.notify_all();
}

// Admits queued requests for one idle slot and launches them; lock held.
void Dispatch(csg_server* s, Slot& sl) {
  std::vector<int64_t> take;
  while (!s->queue.empty() && static_cast<int32_t>(take.size()) < s->cfg.max_batch_jobs) {
    Request& r = s->reqs[s->queue.front()];
    if (s->used + static_cast<int64_t>(r.bytes) > s->outputBudget) break;  // wait for releases
    const Block blk = TakeBlock(s, r.bytes);
    if (!blk.mem) break;
    s->used += static_cast<int64_t>(r.bytes);
    s->stats.peak_bytes_in_use = std::max(s->stats.peak_bytes_in_use, s->used + s->workspace);
    r.block = blk;
    const int64_t px = static_cast<int64_t>(r.req.width) * r.req.height;
    char* b = static_cast<char*>(blk.mem);
    r.out.rgba = reinterpret_cast<float*>(b);
    r.out.depth = reinterpret_cast<float*>(b + Align256(16 * px));
    r.out.prim_id = reinterpret_cast<int32_t*>(b + Align256(16 * px) + Align256(4 * px));
    r.out.flags = reinterpret_cast<uint8_t*>(b + Align256(16 * px) + 2 * Align256(4 * px));
    take.push_back(s->queue.front());
    s->queue.pop_front();
  }
  if (take.empty()) return;
  const double t = Since(s);
  std::vector<csgray::Job> jobs;
  std::vector<int64_t> run;
  jobs.reserve(take.size());
  int64_t base = 0;
  int32_t depth = 1, maxOps = 1, maxLeaves = 1;
  bool anyInBoxes = false;
  csg_render_opts opts{};
  opts.mode = s->cfg.mode;
  bool failed = false;
  for (int64_t tk : take) {
    Request& r = s->reqs[tk];
    const csg_scene* sc = s->scenes[r.req.scene_id];
    csgray::Job jb;
    const csg_status js = BuildJob(sc, &r.req.camera, &r.req.light, &r.req.phong, r.req.width, r.req.height, &r.out,
                                   &opts, s->device, base, &jb);
    r.stats.start_s = t;
    if (js != CSG_OK) {
      // a request that fails validation is not launched; it is done at once
      r.stats.status = js;
      Finish(s, r, js, t);
      failed = true;
      continue;
    }
    // the block's last copy out (a released request's) completes before this batch writes it
    if (r.block.copied) {
      if (hipStreamWaitEvent(sl.stream, r.block.copied, 0) == hipSuccess) {
        sl.waited.push_back(r.block.copied);
      } else {
        (void)hipEventSynchronize(r.block.copied);
        ReturnEvent(s, r.block.copied);
      }
      r.block.copied = nullptr;
    }
    jobs.push_back(jb);
    run.push_back(tk);
    base += jb.tiles;
    depth = std::max(depth, sc->dev.opDepth);
    maxOps = std::max(maxOps, sc->dev.nOps);
    maxLeaves = std::max(maxLeaves, sc->dev.opLeaves);
    anyInBoxes |= sc->dev.conservative != 0;
    r.state = kRunning;
  }
  if (failed) s->doneCv.notify_all();
  if (run.empty()) return;
  for (int64_t tk : run) s->reqs[tk].stats.batch_jobs = static_cast<int32_t>(run.size());
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
    PrepareCoop(s->launch, maxOps, maxLeaves, anyInBoxes, heavy, shape, &p);
    if (csgray::LaunchRender(p, s->launch, shape, sl.stream) != hipSuccess) st = CSG_E_HIP;
  }
  if (st == CSG_OK && hipEventRecord(sl.done, sl.stream) != hipSuccess) st = CSG_E_HIP;
  // A failed launch still retires its requests (outputs undefined, status
  // set) so that no client blocks forever.
  if (st != CSG_OK) (void)hipEventRecord(sl.done, sl.stream);
  for (int64_t tk : run) s->reqs[tk].stats.status = st;
  sl.tickets = std::move(run);
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

void DestroyStreams(csg_server* s) {
  for (Slot& sl : s->slots) {
    if (sl.stream) {
      (void)hipStreamSynchronize(sl.stream);
      (void)ReleaseStreamWorkspace(s->device, sl.stream);
      (void)hipStreamDestroy(sl.stream);
    }
    if (sl.done) (void)hipEventDestroy(sl.done);
    for (hipEvent_t e : sl.waited) (void)hipEventDestroy(e);
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
  s->workspace = static_cast<int64_t>(StreamWorkspaceBytes(s->cfg.max_batch_jobs)) * s->cfg.streams;
  s->outputBudget = s->cfg.budget_bytes - s->workspace;
  if (st == CSG_OK && s->outputBudget <= 0)
    st = Fail(CSG_E_BUDGET, "csg_server_create: the budget does not cover the streams' workspaces (" +
                                std::to_string(s->workspace) + " bytes)");
  s->slots.resize(static_cast<size_t>(s->cfg.streams));
  for (Slot& sl : s->slots) {
    if (st != CSG_OK) break;
    if (hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&sl.done, hipEventDisableTiming) != hipSuccess) {
      st = Fail(CSG_E_HIP, "stream or event creation failed");
      break;
    }
    // the stream's workspace now, inside the budget (not at its first batch)
    uint32_t* q = nullptr;
    uint2* h = nullptr;
    st = QueueFor(s->device, sl.stream, &q);
    if (st == CSG_OK) st = HeavyFor(s->device, sl.stream, &h);
  }
  if (st != CSG_OK) {
    DestroyStreams(s);
    delete s;
    return st;
  }
  s->stats.peak_bytes_in_use = s->workspace;
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
  if (static_cast<int64_t>(bytes) > s->outputBudget) {
    s->stats.rejected++;
    return Fail(CSG_E_BUDGET, "csg_server_submit: the request's outputs exceed the memory budget left beside the "
                              "streams' workspaces");
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
  // Held through the copies: the event recorded after them goes with the
  // block, and a batch that reuses the block waits for it.
  std::lock_guard<std::mutex> lock(s->mu);
  auto it = s->reqs.find(ticket);
  if (it == s->reqs.end()) return Fail(CSG_E_ARG, "csg_server_copy: unknown ticket");
  Request& r = it->second;
  if (r.state != kDone) return Fail(CSG_E_ARG, "csg_server_copy: request not done");
  if (!r.block.mem) return Fail(CSG_E_ARG, "csg_server_copy: the request has no outputs (it failed)");
  const int64_t px = static_cast<int64_t>(r.req.width) * r.req.height;
  const hipStream_t hs = static_cast<hipStream_t>(stream);
  const void* from[4] = {r.out.rgba, r.out.depth, r.out.prim_id, r.out.flags};
  void* to[4] = {dst->rgba, dst->depth, dst->prim_id, dst->flags};
  const int64_t sz[4] = {16 * px, 4 * px, 4 * px, px};
  for (int k = 0; k < 4; k++) {
    if (!to[k]) continue;
    const hipError_t e = hipMemcpyAsync(to[k], from[k], static_cast<size_t>(sz[k]), hipMemcpyDeviceToDevice, hs);
    if (e != hipSuccess) return FailHip(e, "hipMemcpyAsync of request outputs");
  }
  if (!r.block.copied) {
    if (!s->eventPool.empty()) {
      r.block.copied = s->eventPool.back();
      s->eventPool.pop_back();
    } else if (hipEventCreateWithFlags(&r.block.copied, hipEventDisableTiming) != hipSuccess) {
      r.block.copied = nullptr;
      const hipError_t e = hipStreamSynchronize(hs);  // no event: finish the copies here
      return e == hipSuccess ? CSG_OK : FailHip(e, "hipStreamSynchronize");
    }
  }
  const hipError_t e = hipEventRecord(r.block.copied, hs);
  if (e != hipSuccess) return FailHip(e, "hipEventRecord after the copies");
  return CSG_OK;
}

csg_status csg_server_release(csg_server* s, int64_t ticket) {
  if (!s) return Fail(CSG_E_ARG, "csg_server_release: NULL server");
  std::lock_guard<std::mutex> lock(s->mu);
  auto it = s->reqs.find(ticket);
  if (it == s->reqs.end()) return Fail(CSG_E_ARG, "csg_server_release: unknown ticket");
  if (it->second.state != kDone) return Fail(CSG_E_ARG, "csg_server_release: request not done");
  if (it->second.block.mem) {
    s->cache.insert({it->second.bytes, it->second.block});
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
  st->bytes_in_use = s->used + s->workspace;
  st->workspace_bytes = s->workspace;
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
  (void)hipDeviceSynchronize();  // copies out of the blocks on callers' streams
  DestroyStreams(s);
  for (auto& kv : s->reqs) {
    if (kv.second.block.mem) (void)hipFree(kv.second.block.mem);
    if (kv.second.block.copied) (void)hipEventDestroy(kv.second.block.copied);
  }
  for (auto& kv : s->cache) {
    (void)hipFree(kv.second.mem);
    if (kv.second.copied) (void)hipEventDestroy(kv.second.copied);
  }
  for (hipEvent_t e : s->eventPool) (void)hipEventDestroy(e);
  delete s;
  return CSG_OK;
}

}  // extern "C"
