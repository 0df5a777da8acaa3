// Captured batches: csg_batch_create / update / launch / autotune / free.
//
// A batch owns, per launch of at most kMaxBatchJobs jobs, a device job table,
// a pinned staging copy of it and a tile queue.  Its launches are captured
// once into a hipGraph.  An update rewrites the job tables in stream order and
// patches the kernel nodes whose launch parameters changed; a launch replays
// the graph.
//
// Depth classes: with a split depth D > 0 the jobs whose trees have at most D
// operator levels and the deeper ones go to separate launches, so that each
// launch's LDS stacks hold only as many slots as its own deepest tree needs
// (one 8-byte slot per level and lane) and the shallow trees' workgroups are
// not sized for the deep ones.  The autotuner tries no split and one split
// (HalfWorkDepth).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <new>
#include <vector>

#include "csg.h"
#include "device.hpp"
#include "runtime.hpp"

using namespace csgray::rt;

namespace {

struct Chunk {
  csgray::Job* devJobs = nullptr;
  csgray::Job* hostJobs = nullptr;  // pinned
  int32_t cap = 0;                  // jobs the table holds
  uint32_t* queue = nullptr;
  uint2* heavy = nullptr;
  csgray::RenderParams params{};
  csgray::LaunchShape shape{};
  hipGraphNode_t node = nullptr;
  hipGraphNode_t coopNode = nullptr;
};

}  // namespace

struct csg_batch {
  int device = 0;
  csg_render_opts opts{};
  csgray::LaunchConfig cfg;
  bool heaviestFirst = false;
  int32_t depthSplit = 0;  // 0: one class; else jobs of depth <= depthSplit launch apart from deeper ones
  std::vector<csg_job> jobs;
  std::vector<Chunk> chunks;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipEvent_t staged = nullptr;  // last staging copy
};

namespace {

csg_status AllocChunk(Chunk& c, int32_t cap) {
  hipError_t e = hipMalloc(&c.devJobs, sizeof(csgray::Job) * cap);
  if (e == hipSuccess) e = hipHostMalloc(&c.hostJobs, sizeof(csgray::Job) * cap, hipHostMallocDefault);
  if (e == hipSuccess) e = hipMalloc(&c.queue, csgray::kQueueBytes);
  if (e == hipSuccess) e = hipMemset(c.queue, 0, csgray::kQueueBytes);
  if (e == hipSuccess) e = hipMalloc(&c.heavy, csgray::kHeavyBytes);
  if (e != hipSuccess) return e == hipErrorOutOfMemory ? Fail(CSG_E_OOM, "batch allocation failed") : FailHip(e, "batch allocation");
  c.cap = cap;
  return CSG_OK;
}

void FreeChunks(csg_batch* b) {
  for (Chunk& c : b->chunks) {
    if (c.devJobs) (void)hipFree(c.devJobs);
    if (c.hostJobs) (void)hipHostFree(c.hostJobs);
    if (c.queue) (void)hipFree(c.queue);
    if (c.heavy) (void)hipFree(c.heavy);
  }
  b->chunks.clear();
}

void DestroyGraph(csg_batch* b) {
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
}

// The batch's launches: job indices in launch order, and for each launch
// (chunk) its first position in that order and its job count.  Jobs go in
// the given order or heaviest first, then (with a depth split) the shallow
// class before the deep one; each class is cut into chunks of at most
// kMaxBatchJobs jobs.
struct Plan {
  std::vector<int32_t> order;
  std::vector<std::pair<int32_t, int32_t>> chunks;  // (start, count)
};

csg_status MakePlan(const csg_batch* b, Plan* plan) {
  const int32_t n = static_cast<int32_t>(b->jobs.size());
  plan->order.resize(static_cast<size_t>(n));
  for (int32_t i = 0; i < n; i++) plan->order[i] = i;
  if (b->heaviestFirst) {
    const csg_status st = csg_batch_order(b->jobs.data(), n, plan->order.data());
    if (st != CSG_OK) return st;
  }
  int32_t shallow = n;
  if (b->depthSplit > 0) {
    auto mid = std::stable_partition(plan->order.begin(), plan->order.end(), [&](int32_t i) {
      return b->jobs[i].scene && b->jobs[i].scene->dev.opDepth <= b->depthSplit;
    });
    shallow = static_cast<int32_t>(mid - plan->order.begin());
  }
  plan->chunks.clear();
  for (const auto& cls : {std::make_pair(0, shallow), std::make_pair(shallow, n)})
    for (int32_t s = cls.first; s < cls.second; s += csgray::kMaxBatchJobs)
      plan->chunks.push_back({s, std::min(cls.second - s, csgray::kMaxBatchJobs)});
  return CSG_OK;
}

// Allocates chunks until there are as many as the batch's launch plan has.
csg_status EnsureChunks(csg_batch* b) {
  Plan plan;
  const csg_status st = MakePlan(b, &plan);
  if (st != CSG_OK) return st;
  while (b->chunks.size() < plan.chunks.size()) {
    b->chunks.emplace_back();
    const csg_status a = AllocChunk(b->chunks.back(), csgray::kMaxBatchJobs);
    if (a != CSG_OK) return a;
  }
  return CSG_OK;
}

// Frees the chunks beyond the launch plan (the device is idle).
void TrimChunks(csg_batch* b) {
  Plan plan;
  if (MakePlan(b, &plan) != CSG_OK) return;
  const size_t keep = std::max<size_t>(plan.chunks.size(), 1);
  while (b->chunks.size() > keep) {
    Chunk& c = b->chunks.back();
    if (c.devJobs) (void)hipFree(c.devJobs);
    if (c.hostJobs) (void)hipHostFree(c.hostJobs);
    if (c.queue) (void)hipFree(c.queue);
    if (c.heavy) (void)hipFree(c.heavy);
    b->chunks.pop_back();
  }
}

// The split depth the autotuner tries: the smallest d with at least half of
// the batch's work (pixels x (1 + depth)) in trees of depth <= d, or, when
// that is the deepest tree, the largest depth below it; 0 when all trees
// have one depth (no second class).
int32_t HalfWorkDepth(const csg_batch* b) {
  std::vector<std::pair<int32_t, double>> w;
  double total = 0;
  for (const csg_job& j : b->jobs) {
    const int32_t d = j.scene ? j.scene->dev.opDepth : 0;
    const double work = static_cast<double>(j.width) * j.height * (1.0 + d);
    w.push_back({d, work});
    total += work;
  }
  std::sort(w.begin(), w.end());
  if (w.empty() || w.front().first == w.back().first) return 0;
  double acc = 0;
  for (const auto& x : w) {
    acc += x.second;
    if (acc >= 0.5 * total && x.first < w.back().first) return x.first;
  }
  int32_t below = 0;
  for (const auto& x : w)
    if (x.first < w.back().first) below = x.first;
  return below;
}

// Fills each chunk's staging table and launch parameters from b->jobs in the
// batch's launch plan (chunks beyond the plan get no jobs).  Does not touch
// device memory.
csg_status Stage(csg_batch* b) {
  Plan plan;
  const csg_status pst = MakePlan(b, &plan);
  if (pst != CSG_OK) return pst;
  if (plan.chunks.size() > b->chunks.size())
    return Fail(CSG_E_ARG, "csg_batch: more jobs than the batch was created for");
  for (int32_t c = 0; c < static_cast<int32_t>(b->chunks.size()); c++) {
    Chunk& ch = b->chunks[c];
    const int32_t start = c < static_cast<int32_t>(plan.chunks.size()) ? plan.chunks[c].first : 0;
    const int32_t cnt = c < static_cast<int32_t>(plan.chunks.size()) ? plan.chunks[c].second : 0;
    if (cnt > ch.cap) return Fail(CSG_E_ARG, "csg_batch: chunk capacity exceeded");
    int64_t base = 0;
    int32_t depth = 1, maxOps = 1, maxLeaves = 1;
    bool anyInBoxes = false;
    for (int32_t i = 0; i < cnt; i++) {
      const csg_job& j = b->jobs[plan.order[start + i]];
      maxOps = std::max(maxOps, j.scene->dev.nOps);
      maxLeaves = std::max(maxLeaves, j.scene->dev.opLeaves);
      anyInBoxes |= j.scene->dev.conservative != 0;
      const csg_status st = BuildJob(j.scene, &j.camera, &j.light, &j.phong, j.width, j.height, &j.out, &b->opts,
                                     b->device, base, &ch.hostJobs[i]);
      if (st != CSG_OK) return st;
      base += ch.hostJobs[i].tiles;
      depth = std::max(depth, j.scene->dev.opDepth);
    }
    csgray::RenderParams& p = ch.params;
    std::memset(&p, 0, sizeof(p));
    p.jobs = ch.devJobs;
    p.numJobs = std::max(cnt, 1);
    p.numTiles = static_cast<int32_t>(base);  // an empty chunk has no tiles: every wave exits at once
    p.stackSlots = depth;
    p.queue = ch.queue;
    p.morton = b->cfg.morton ? 1 : 0;
    if (cnt == 0) std::memset(&ch.hostJobs[0], 0, sizeof(csgray::Job));
    const hipError_t e = csgray::ChooseShape(depth, 0, b->cfg, &ch.shape, std::max(cnt, 1));
    if (e != hipSuccess) return FailHip(e, "choosing the launch shape");
    csgray::FitGridToTiles(&ch.shape, base);
    p.totalWaves = static_cast<uint32_t>(ch.shape.grid) * static_cast<uint32_t>(ch.shape.block / 64);
    PrepareCoop(b->cfg, maxOps, maxLeaves, anyInBoxes, ch.heavy, ch.shape, &p);
  }
  return CSG_OK;
}

// Copies every chunk's staging table to the device in stream order.
csg_status Upload(csg_batch* b, hipStream_t stream) {
  for (Chunk& ch : b->chunks) {
    const hipError_t e = hipMemcpyAsync(ch.devJobs, ch.hostJobs, sizeof(csgray::Job) * std::max(ch.params.numJobs, 1),
                                        hipMemcpyHostToDevice, stream);
    if (e != hipSuccess) return FailHip(e, "hipMemcpyAsync of a batch job table");
  }
  const hipError_t e = hipEventRecord(b->staged, stream);
  if (e != hipSuccess) return FailHip(e, "hipEventRecord");
  return CSG_OK;
}

// Waits until the staging tables may be rewritten.
csg_status WaitStaged(csg_batch* b) {
  const hipError_t e = hipEventSynchronize(b->staged);
  if (e != hipSuccess) return FailHip(e, "hipEventSynchronize");
  return CSG_OK;
}

csg_status LaunchDirect(csg_batch* b, hipStream_t stream) {
  for (Chunk& ch : b->chunks) {
    const hipError_t e = csgray::LaunchRender(ch.params, b->cfg, ch.shape, stream);
    if (e != hipSuccess) return FailHip(e, "batched render launch");
  }
  return CSG_OK;
}

// Captures the chunks' launches into a graph and finds each kernel node.
csg_status Capture(csg_batch* b) {
  DestroyGraph(b);
  hipStream_t cs = nullptr;
  hipError_t e = hipStreamCreateWithFlags(&cs, hipStreamNonBlocking);
  if (e != hipSuccess) return FailHip(e, "hipStreamCreate");
  e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
  csg_status st = CSG_OK;
  if (e == hipSuccess) {
    st = LaunchDirect(b, cs);
    const hipError_t e2 = hipStreamEndCapture(cs, &b->graph);
    if (st == CSG_OK && e2 != hipSuccess) st = FailHip(e2, "hipStreamEndCapture");
  } else {
    st = FailHip(e, "hipStreamBeginCapture");
  }
  (void)hipStreamDestroy(cs);
  if (st != CSG_OK) return st;
  e = hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0);
  if (e != hipSuccess) return FailHip(e, "hipGraphInstantiate");
  size_t count = 0;
  e = hipGraphGetNodes(b->graph, nullptr, &count);
  if (e != hipSuccess) return FailHip(e, "hipGraphGetNodes");
  std::vector<hipGraphNode_t> nodes(count);
  e = hipGraphGetNodes(b->graph, nodes.data(), &count);
  if (e != hipSuccess) return FailHip(e, "hipGraphGetNodes");
  for (Chunk& ch : b->chunks) {
    ch.node = nullptr;
    ch.coopNode = nullptr;
  }
  for (hipGraphNode_t nd : nodes) {
    hipGraphNodeType type;
    if (hipGraphNodeGetType(nd, &type) != hipSuccess || type != hipGraphNodeTypeKernel) continue;
    hipKernelNodeParams kp{};
    if (hipGraphKernelNodeGetParams(nd, &kp) != hipSuccess || !kp.kernelParams) continue;
    const auto* p = static_cast<const csgray::RenderParams*>(kp.kernelParams[0]);
    for (Chunk& ch : b->chunks) {
      if (p->queue != ch.queue) continue;
      // the render kernel runs on the chunk's launch shape, the cooperative one on 64-lane groups
      if (kp.blockDim.x == 64 && ch.shape.block != 64) ch.coopNode = nd;
      else ch.node = nd;
    }
  }
  for (const Chunk& ch : b->chunks) {
    if (!ch.node) return Fail(CSG_E_HIP, "captured graph lacks a render node");
    if ((ch.params.budget > 0) != (ch.coopNode != nullptr)) return Fail(CSG_E_HIP, "captured graph lacks a node");
  }
  return CSG_OK;
}

// Points each kernel node of the instantiated graph at its chunk's current
// launch parameters.  The graph is captured again instead when the nodes no
// longer match the chunks (another chunk count, deferral switched on or off)
// or a kernel needs more than 64 KB of LDS, whose launch attribute only a
// direct launch sets.
csg_status PatchNodes(csg_batch* b) {
  for (const Chunk& ch : b->chunks) {
    const size_t coopLds = csgray::CoopLaunchLds(ch.params);
    if (!ch.node || (ch.params.budget > 0) != (ch.coopNode != nullptr) || ch.shape.lds > 64 * 1024 ||
        (ch.params.budget > 0 && coopLds > 64 * 1024))
      return Capture(b);
  }
  for (Chunk& ch : b->chunks) {
    hipKernelNodeParams kp{};
    hipError_t e = hipGraphKernelNodeGetParams(ch.node, &kp);
    if (e != hipSuccess) return FailHip(e, "hipGraphKernelNodeGetParams");
    void* args[1] = {&ch.params};
    kp.kernelParams = args;
    kp.extra = nullptr;
    kp.gridDim = dim3(ch.shape.grid);
    kp.blockDim = dim3(ch.shape.block);
    kp.sharedMemBytes = static_cast<unsigned int>(ch.shape.lds);
    e = hipGraphExecKernelNodeSetParams(b->exec, ch.node, &kp);
    if (e != hipSuccess) return FailHip(e, "hipGraphExecKernelNodeSetParams");
    if (ch.coopNode) {
      hipKernelNodeParams cp{};
      e = hipGraphKernelNodeGetParams(ch.coopNode, &cp);
      if (e != hipSuccess) return FailHip(e, "hipGraphKernelNodeGetParams");
      void* cargs[1] = {&ch.params};
      cp.kernelParams = cargs;
      cp.extra = nullptr;
      cp.gridDim = dim3(ch.params.coopWaves);
      cp.blockDim = dim3(64);
      cp.sharedMemBytes = static_cast<unsigned int>(csgray::CoopLaunchLds(ch.params));
      e = hipGraphExecKernelNodeSetParams(b->exec, ch.coopNode, &cp);
      if (e != hipSuccess) return FailHip(e, "hipGraphExecKernelNodeSetParams");
    }
  }
  return CSG_OK;
}

}  // namespace

extern "C" {

csg_status csg_batch_create(const csg_job* jobs, int32_t njobs, const csg_render_opts* opts, csg_batch** out) {
  if (!out) return Fail(CSG_E_ARG, "csg_batch_create: out is NULL");
  *out = nullptr;
  if (!jobs || njobs <= 0) return Fail(CSG_E_ARG, "csg_batch_create: no jobs");
  if (opts && opts->tile_order) return Fail(CSG_E_ARG, "csg_batch_create: tile_order applies to single renders");
  if (opts && opts->counters) return Fail(CSG_E_ARG, "csg_batch_create: counters are not supported in batches");
  csg_batch* b = new (std::nothrow) csg_batch();
  if (!b) return Fail(CSG_E_OOM, "csg_batch_create: host allocation failed");
  if (opts) b->opts = *opts;
  b->opts.scene_placement = 1;
  csg_status st = ParseOpts(&b->opts, &b->cfg);
  hipError_t e = hipSuccess;
  if (st == CSG_OK) {
    e = hipGetDevice(&b->device);
    if (e != hipSuccess) st = FailHip(e, "hipGetDevice");
  }
  if (st == CSG_OK) {
    e = hipEventCreateWithFlags(&b->staged, hipEventDisableTiming);
    if (e != hipSuccess) st = FailHip(e, "hipEventCreate");
  }
  if (st == CSG_OK) {
    b->jobs.assign(jobs, jobs + njobs);
    const int32_t nch = (njobs + csgray::kMaxBatchJobs - 1) / csgray::kMaxBatchJobs;
    b->chunks.resize(static_cast<size_t>(nch));
    for (int32_t c = 0; c < nch && st == CSG_OK; c++) st = AllocChunk(b->chunks[c], csgray::kMaxBatchJobs);
  }
  if (st == CSG_OK) st = Stage(b);
  if (st == CSG_OK) st = Upload(b, nullptr);
  if (st == CSG_OK) {
    e = hipStreamSynchronize(nullptr);
    if (e != hipSuccess) st = FailHip(e, "hipStreamSynchronize");
  }
  if (st == CSG_OK) st = Capture(b);
  if (st != CSG_OK) {
    const std::string keep = csg_last_error();
    (void)csg_batch_free(b);
    return Fail(st, keep);
  }
  *out = b;
  return CSG_OK;
}

csg_status csg_batch_update(csg_batch* b, const csg_job* jobs, int32_t njobs, void* stream) {
  if (!b || !jobs || njobs <= 0) return Fail(CSG_E_ARG, "csg_batch_update: bad argument");
  csg_status st = WaitStaged(b);
  if (st != CSG_OK) return st;
  std::vector<csg_job> keep = b->jobs;
  b->jobs.assign(jobs, jobs + njobs);
  st = EnsureChunks(b);
  if (st == CSG_OK) st = Stage(b);
  if (st != CSG_OK) {
    b->jobs = keep;
    (void)Stage(b);
    return st;
  }
  st = Upload(b, static_cast<hipStream_t>(stream));
  if (st != CSG_OK) return st;
  return PatchNodes(b);
}

csg_status csg_batch_launch(csg_batch* b, void* stream) {
  if (!b || !b->exec) return Fail(CSG_E_ARG, "csg_batch_launch: no batch");
  const hipError_t e = hipGraphLaunch(b->exec, static_cast<hipStream_t>(stream));
  if (e != hipSuccess) return FailHip(e, "hipGraphLaunch");
  return CSG_OK;
}

csg_status csg_batch_autotune(csg_batch* b, int32_t reps, void* stream, csg_batch_tuning* chosen) {
  if (!b) return Fail(CSG_E_ARG, "csg_batch_autotune: no batch");
  const hipStream_t hs = static_cast<hipStream_t>(stream);
  reps = std::max(1, reps);
  hipEvent_t t0, t1;
  if (hipEventCreate(&t0) != hipSuccess || hipEventCreate(&t1) != hipSuccess)
    return Fail(CSG_E_HIP, "hipEventCreate failed");
  const csg_render_opts base = b->opts;
  const bool baseHeavy = b->heaviestFirst;
  const int32_t baseSplit = b->depthSplit;
  float best = 1e30f;
  csg_render_opts bestOpts = base;
  bool bestHeavy = baseHeavy;
  int32_t bestSplit = baseSplit;
  csg_status st = CSG_OK;
  std::vector<int32_t> splits{0};
  if (const int32_t half = HalfWorkDepth(b)) splits.push_back(half);
  for (int32_t split : splits)
  for (int32_t block : {256, 512, 1024}) {
    for (int32_t walk : {0, 1}) {
      for (bool heavy : {false, true}) {
        csg_render_opts o = base;
        o.block = block;
        o.tile_walk = walk;
        b->opts = o;
        b->heaviestFirst = heavy;
        b->depthSplit = split;
        st = ParseOpts(&b->opts, &b->cfg);
        if (st == CSG_OK) st = WaitStaged(b);
        if (st == CSG_OK) st = EnsureChunks(b);
        if (st == CSG_OK) st = Stage(b);
        if (st == CSG_OK) st = Upload(b, hs);
        if (st == CSG_OK) st = LaunchDirect(b, hs);  // warm-up
        if (st != CSG_OK) break;
        (void)hipEventRecord(t0, hs);
        for (int32_t r = 0; r < reps && st == CSG_OK; r++) st = LaunchDirect(b, hs);
        (void)hipEventRecord(t1, hs);
        if (hipEventSynchronize(t1) != hipSuccess) st = Fail(CSG_E_HIP, "hipEventSynchronize failed");
        if (st != CSG_OK) break;
        float ms = 0;
        (void)hipEventElapsedTime(&ms, t0, t1);
        ms /= static_cast<float>(reps);
        if (ms < best) {
          best = ms;
          bestOpts = o;
          bestHeavy = heavy;
          bestSplit = split;
        }
      }
      if (st != CSG_OK) break;
    }
    if (st != CSG_OK) break;
  }
  (void)hipEventDestroy(t0);
  (void)hipEventDestroy(t1);
  b->opts = st == CSG_OK ? bestOpts : base;
  b->heaviestFirst = st == CSG_OK ? bestHeavy : baseHeavy;
  b->depthSplit = st == CSG_OK ? bestSplit : baseSplit;
  csg_status st2 = ParseOpts(&b->opts, &b->cfg);
  if (st2 == CSG_OK) st2 = WaitStaged(b);
  if (st2 == CSG_OK && hipStreamSynchronize(hs) != hipSuccess) st2 = Fail(CSG_E_HIP, "hipStreamSynchronize failed");
  if (st2 == CSG_OK) {
    DestroyGraph(b);  // the graph may point at chunks about to be freed
    TrimChunks(b);
  }
  if (st2 == CSG_OK) st2 = Stage(b);
  if (st2 == CSG_OK) st2 = Upload(b, hs);
  if (st2 == CSG_OK && hipStreamSynchronize(hs) != hipSuccess) st2 = Fail(CSG_E_HIP, "hipStreamSynchronize failed");
  if (st2 == CSG_OK) st2 = Capture(b);
  if (st != CSG_OK) return st;
  if (st2 != CSG_OK) return st2;
  if (chosen) {
    chosen->block = b->opts.block;
    chosen->tile_walk = b->opts.tile_walk;
    chosen->heaviest_first = b->heaviestFirst ? 1 : 0;
    chosen->ms = best;
    chosen->depth_split = b->depthSplit;
  }
  return CSG_OK;
}

csg_status csg_batch_free(csg_batch* b) {
  if (!b) return CSG_OK;
  if (b->staged) {
    (void)hipEventSynchronize(b->staged);
    (void)hipEventDestroy(b->staged);
  }
  (void)hipDeviceSynchronize();
  DestroyGraph(b);
  FreeChunks(b);
  delete b;
  return CSG_OK;
}

}  // extern "C"
