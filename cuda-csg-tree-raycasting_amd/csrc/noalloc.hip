// csg_render_noalloc: csg_render_ex without any device allocation, for
// callers that cannot allocate (code under stream capture, tight serving
// loops).  It has no counterpart in the reference.
//
// csg_render_ex takes its scratch (the tile and heavy-ray queue words, the
// heavy-ray list and the resume area of handed-over rays) from a workspace
// that the library allocates per (device, stream) on the first render.  Here
// the same scratch lives in one static device array per GPU, `scratch`
// (kScratchBytes, about 48 MiB, held by each GPU the library runs on until
// the process ends):
//   [0, kQueueRegion)                     queue words (zero between launches)
//   [kQueueRegion, + 8 B x kHeavyCap)     heavy-ray list, one entry per ray
//   [..., + kNoAllocResumeBytes)          resume area (16 MiB)
// The list has room for every ray of a call of up to kNoAllocMaxRays rays,
// and larger calls are rejected.  The resume area holds the state of the
// first rays handed over (about 73,000 with 10 stack slots) and, in its last
// eighth, the directions of the first 131,072; later ones restart from the
// root, or are built again from their pixel, in the cooperative kernel, with
// the same results.
//
// All calls on one GPU share its array: calls that overlap on one GPU (two
// streams, two host threads) race.  The kernels return the queue words to
// zero at the end of every launch, and the set-up clears them once.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>

#include "csg.h"
#include "device.hpp"
#include "runtime.hpp"

namespace {

constexpr size_t kQueueRegion = (csgray::kQueueBytes + 255) & ~static_cast<size_t>(255);
constexpr size_t kNoAllocResumeBytes = size_t{16} << 20;
constexpr size_t kScratchBytes = kQueueRegion + sizeof(uint2) * csgray::kHeavyCap + kNoAllocResumeBytes;
constexpr int64_t kNoAllocMaxRays = int64_t{csgray::kHeavyCap};
constexpr int kMaxDevices = 64;

}  // namespace

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

__global__ void ClearWords(uint32_t* w, uint32_t n) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) w[i] = 0u;
}

std::mutex g_scratchMutex;
unsigned char* g_scratch[kMaxDevices] = {};

// The device address of `scratch` on device `dev` (current), found once per
// device; the first call also clears the queue words on `stream` (a kernel,
// not a memset, so that a capturing caller records an ordinary kernel node).
csg_status ScratchFor(int dev, hipStream_t stream, unsigned char** out) {
  using csgray::rt::Fail;
  using csgray::rt::FailHip;
  if (dev < 0 || dev >= kMaxDevices) return Fail(CSG_E_ARG, "csg_render_noalloc: device index out of range");
  std::lock_guard<std::mutex> lock(g_scratchMutex);
  if (!g_scratch[dev]) {
    void* ptr = nullptr;
    hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return FailHip(e, "hipGetSymbolAddress of the no-allocation scratch");
    if (!ptr) return Fail(CSG_E_HIP, "hipGetSymbolAddress returned no address for the scratch");
    hipLaunchKernelGGL(ClearWords, dim3(16), dim3(256), 0, stream, static_cast<uint32_t*>(ptr),
                       static_cast<uint32_t>(kQueueRegion / 4));
    e = hipGetLastError();
    if (e != hipSuccess) return FailHip(e, "clearing the no-allocation scratch");
    g_scratch[dev] = static_cast<unsigned char*>(ptr);
  }
  *out = g_scratch[dev];
  return CSG_OK;
}

}  // namespace

using namespace csgray::rt;

extern "C" csg_status csg_render_noalloc(const csg_scene* s, const csg_camera* cam, const csg_light* light,
                                         const csg_phong* phong, int32_t width, int32_t height,
                                         const csg_outputs* out, const csg_render_opts* opts, void* stream) {
  csgray::LaunchConfig cfg;
  const csg_status ost = ParseOpts(opts, &cfg);
  if (ost != CSG_OK) return ost;
  if (!s || !cam || !light || !phong || !out) return Fail(CSG_E_ARG, "csg_render_noalloc: NULL argument");
  if (!s->dmem) return Fail(CSG_E_ARG, "csg_render_noalloc: scene is not on a device (call csg_scene_upload)");
  const int64_t spp = opts && opts->samples > 1 ? opts->samples : 1;
  if (width >= 2 && height >= 2 && static_cast<int64_t>(width) * height * spp > kNoAllocMaxRays)
    return Fail(CSG_E_ARG, "csg_render_noalloc: at most " + std::to_string(kNoAllocMaxRays) +
                               " rays per call (its static heavy-ray list holds one entry per ray); "
                               "use csg_render_ex for larger frames");
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return FailHip(e, "hipGetDevice");
  csgray::RenderParams p{};
  const csg_status js = BuildJob(s, cam, light, phong, width, height, out, opts, dev, 0, &p.job);
  if (js != CSG_OK) return js;
  p.numJobs = 0;
  p.numTiles = p.job.tiles;
  p.stackSlots = s->dev.opDepth;
  p.tileOrder = opts ? opts->tile_order : nullptr;
  p.morton = cfg.morton ? 1 : 0;
  csgray::LaunchShape shape;
  e = csgray::ChooseShape(s->dev.opDepth, SceneBytes(s), cfg, &shape, 0, p.numTiles, s->dev.nOps);
  if (e != hipSuccess) return FailHip(e, "choosing the launch shape");
  csgray::FitGridToTiles(&shape, p.numTiles);
  p.totalWaves = static_cast<uint32_t>(shape.grid) * static_cast<uint32_t>(shape.block / 64);
  p.cacheOps = shape.cacheOps;
  const hipStream_t hs = static_cast<hipStream_t>(stream);
  unsigned char* base = nullptr;
  const csg_status ss = ScratchFor(dev, hs, &base);
  if (ss != CSG_OK) return ss;
  p.queue = reinterpret_cast<uint32_t*>(base);
  uint2* heavy = reinterpret_cast<uint2*>(base + kQueueRegion);
  PrepareCoop(cfg, s->dev.nOps, s->dev.opLeaves, s->dev.conservative != 0, cfg.budget >= 0 ? heavy : nullptr, shape,
              &p);
  p.resumeBytes = static_cast<uint32_t>(kNoAllocResumeBytes);
  e = csgray::LaunchRender(p, cfg, shape, hs);
  if (e != hipSuccess) return FailHip(e, "render launch");
  return CSG_OK;
}
