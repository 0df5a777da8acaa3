// C ABI (include/csg.h) over the host scene model and the HIP render kernel.
#include "csg.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <map>
#include <mutex>
#include <cstring>
#include <new>
#include <string>

#include "camera.hpp"
#include "device.hpp"
#include "runtime.hpp"
#include "scene.hpp"

namespace csgray {
namespace rt {

thread_local std::string g_lastError;

csg_status Fail(csg_status s, const std::string& msg) {
  g_lastError = msg;
  return s;
}

csg_status FailHip(hipError_t e, const char* what) {
  return Fail(CSG_E_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

void CopyErr(char* err, size_t errlen, const std::string& msg) {
  if (err && errlen) std::snprintf(err, errlen, "%s", msg.c_str());
}

size_t Align256(size_t n) { return (n + 255) & ~static_cast<size_t>(255); }

// One device allocation holding the scene arrays back to back.
csg_status Upload(csg_scene* s) {
  const csgray::FlatScene& f = s->flat;
  const std::vector<csgray::F4>* src[6] = {&f.ops, &f.primA, &f.primB, &f.primC, &f.ivOps, &f.walkOps};
  size_t off[6], total = 0;
  for (int k = 0; k < 6; k++) {
    off[k] = total;
    total += Align256(sizeof(csgray::F4) * src[k]->size());
  }
  s->bytes = total;
  hipError_t e = hipGetDevice(&s->device);
  if (e != hipSuccess) return FailHip(e, "hipGetDevice");
  e = hipMalloc(&s->dmem, std::max<size_t>(s->bytes, 256));
  if (e != hipSuccess) {
    s->dmem = nullptr;
    return e == hipErrorOutOfMemory ? Fail(CSG_E_OOM, "hipMalloc of the scene failed") : FailHip(e, "hipMalloc");
  }
  char* base = static_cast<char*>(s->dmem);
  for (int k = 0; k < 6; k++) {
    if (src[k]->empty()) continue;
    e = hipMemcpy(base + off[k], src[k]->data(), sizeof(csgray::F4) * src[k]->size(), hipMemcpyHostToDevice);
    if (e != hipSuccess) return FailHip(e, "hipMemcpy of the scene");
  }
  s->dev.ops = reinterpret_cast<const float4*>(base + off[0]);
  s->dev.primA = reinterpret_cast<const float4*>(base + off[1]);
  s->dev.primB = reinterpret_cast<const float4*>(base + off[2]);
  s->dev.primC = reinterpret_cast<const float4*>(base + off[3]);
  s->dev.ivOps = f.ivOps.empty() ? nullptr : reinterpret_cast<const float4*>(base + off[4]);
  s->dev.walkOps = f.walkOps.empty() ? nullptr : reinterpret_cast<const float4*>(base + off[5]);
  s->dev.rootDesc = f.rootDesc;
  s->dev.nOps = f.nOps;
  s->dev.nPrims = f.nPrims;
  s->dev.opDepth = f.opDepth;
  s->dev.opLeaves = f.opLeaves;
  s->dev.conservative = f.conservative ? 1u : 0u;
  return CSG_OK;
}

// The render kernels' work queues (tile and heavy-ray counters, time stamps;
// kQueueBytes) per (device, stream), zero between launches (the last
// workgroup of each launch zeroes them).  Renders on one stream are ordered,
// so they can share it; separate streams get separate queues.  A stream's
// queue, heavy-ray list and job table live until ReleaseStreamWorkspace (the
// server's streams) or the process ends.
std::mutex g_queueMutex;
std::map<std::pair<int, hipStream_t>, uint32_t*> g_queues;

csg_status QueueFor(int dev, hipStream_t stream, uint32_t** out) {
  std::lock_guard<std::mutex> lock(g_queueMutex);
  auto it = g_queues.find({dev, stream});
  if (it != g_queues.end()) {
    *out = it->second;
    return CSG_OK;
  }
  void* q = nullptr;
  hipError_t e = hipMalloc(&q, csgray::kQueueBytes);
  if (e != hipSuccess) return e == hipErrorOutOfMemory ? Fail(CSG_E_OOM, "queue allocation failed")
                                                       : FailHip(e, "hipMalloc");
  e = hipMemsetAsync(q, 0, csgray::kQueueBytes, stream);
  if (e != hipSuccess) {
    (void)hipFree(q);
    return FailHip(e, "hipMemsetAsync");
  }
  g_queues[{dev, stream}] = static_cast<uint32_t*>(q);
  *out = static_cast<uint32_t*>(q);
  return CSG_OK;
}

std::mutex g_heavyMutex;
std::map<std::pair<int, hipStream_t>, uint2*> g_heavy;

csg_status HeavyFor(int dev, hipStream_t stream, uint2** out) {
  std::lock_guard<std::mutex> lock(g_heavyMutex);
  auto it = g_heavy.find({dev, stream});
  if (it != g_heavy.end()) {
    *out = it->second;
    return CSG_OK;
  }
  void* h = nullptr;
  const hipError_t e = hipMalloc(&h, csgray::kHeavyBytes);
  if (e != hipSuccess) return e == hipErrorOutOfMemory ? Fail(CSG_E_OOM, "heavy-ray list allocation failed")
                                                       : FailHip(e, "hipMalloc");
  g_heavy[{dev, stream}] = static_cast<uint2*>(h);
  *out = static_cast<uint2*>(h);
  return CSG_OK;
}

// Budget (0: no deferral), grid and LDS of the cooperative kernel for scenes
// of at most `maxOps` operators and `maxLeaves` operators without an
// operator child: one-wave workgroups over the scene in global memory, up to
// 16 per CU as their LDS allows.
struct CoopPlan {
  int32_t budget = 0;
  uint32_t waves = 0;
  uint32_t queueCap = 0;
  size_t ldsBytes = 0;
};

// Interval classification applies to fast-mode launches without counters
// that hold a scene whose primitives all lie in their boxes.
bool CoopLists(const LaunchConfig& cfg, bool anyInBoxes) {
  return cfg.coopIntervals && cfg.mode == Mode::kFast && !cfg.count && anyInBoxes;
}

CoopPlan PlanCoop(const LaunchConfig& cfg, int32_t maxOps, int32_t maxLeaves, bool lists) {
  CoopPlan c;
  if (cfg.budget < 0 || maxOps <= 0) return c;
  const size_t wave = CoopLdsBytes(maxOps, maxLeaves, lists, &c.queueCap);
  if (wave == 0) return c;
  int cus = 0;
  size_t ldsPerBlock = 0, ldsPerCu = 0;
  if (DeviceShapeLimits(&cus, &ldsPerBlock, &ldsPerCu) != hipSuccess || cus <= 0) return c;
  const size_t perCu = std::max<size_t>(1, std::min<size_t>(16, ldsPerCu / (wave + 256)));
  c.budget = cfg.budget > 0 ? cfg.budget : (lists ? kDefaultIntervalBudget : kDefaultBudget);
  c.waves = static_cast<uint32_t>(static_cast<size_t>(cus) * perCu);
  c.ldsBytes = wave;
  return c;
}

// Default tail grace, per mille of the time the tiles took to hand out.
// After the last tile is taken, the tiles in flight finish in about one
// tile's time, while handing out took about tiles / waves tile times: the
// grace is k x waves / tiles per mille, k = 9500 for single-scene launches
// (950 % x waves / tiles) and 5500 for batched ones (550 %).  Measured
// best: 120-130 % for 1024-lane workgroups on the 1920x1080 cheese frame
// (k = 9500 gives 120 %), about 3 % for the 256-tree batch (k = 5500).
// With few tiles a wave the tile hand-out time is no yardstick for the
// tail (the last tile goes out after one or two tile times), so a single
// scene's grace stays at or below 300 %: the 256 random trees of the
// batched mix, each in its own launch, 222 -> 215.5 ms, the cheese frames
// unchanged (a cap of 150 % would cost its 1024x1024 frame 10 %).
// Where the cooperative kernel classifies rays by intervals (`lists`) a
// handed-over ray costs it a fraction of what the per-lane kernel's tail
// costs, so the tail is handed over at once unless the frame has only a few
// tiles a wave (cheese, medians of three interleaved runs, budget 64:
// 3840x2160 and 2560x1440 best with no grace, 1920x1080 with 0-5 %, against
// 1024x1024 60-150 % and 1280x720 35-90 %; waves / tiles = 0.03, 0.07,
// 0.13, 0.25 and 0.28): 95000 % x (waves / tiles)^5, at most 100 % (0.5,
// 0.5, 3, 93 and 100 % for those frames).
int32_t DefaultCoopGraceMille(int64_t waves, int64_t tiles, bool batched, bool lists) {
  if (tiles <= 0) return 1000;
  if (lists && !batched) {
    const double r = static_cast<double>(waves) / static_cast<double>(tiles);
    const double g = 950000.0 * (r * r) * (r * r) * r;
    return static_cast<int32_t>(std::min(1000.0, std::max(5.0, g + 0.5)));
  }
  const int64_t k = batched ? 5500 : 9500;
  const int64_t g = (k * waves + tiles / 2) / tiles;
  return static_cast<int32_t>(std::min<int64_t>(std::max<int64_t>(g, 5), batched ? 1000000 : 3000));
}

void PrepareCoop(const LaunchConfig& cfg, int32_t maxOps, int32_t maxLeaves, bool anyInBoxes, uint2* heavy,
                 const LaunchShape& shape, RenderParams* p) {
  p->budget = 0;
  if (!heavy) return;
  const bool lists = CoopLists(cfg, anyInBoxes);
  const CoopPlan c = PlanCoop(cfg, maxOps, maxLeaves, lists);
  if (c.budget <= 0) return;
  p->budget = c.budget;
  p->heavy = heavy;
  p->heavyCap = kHeavyCap;
  p->resume = reinterpret_cast<uint32_t*>(heavy + kHeavyCap);
  p->resumeBytes = static_cast<uint32_t>(kResumeBytes);
  p->coopOps = maxOps;
  p->coopLists = lists ? 1 : 0;
  p->coopQueueCap = c.queueCap;
  p->coopWaves = c.waves;
  p->coopEager = cfg.coopEager ? 1 : 0;
  p->coopGrace = cfg.coopGrace > 0 ? cfg.coopGrace * 10
                                   : (cfg.coopGrace < 0 ? 0
                                                        : DefaultCoopGraceMille(p->totalWaves, p->numTiles,
                                                                                p->numJobs > 0, lists));
}

}  // namespace rt
}  // namespace csgray

using namespace csgray::rt;

static_assert(csgray::kGroupFlag == csgray::kWalkGroupFlag && csgray::kGroupChunk == csgray::kWalkChunk,
              "the walk records' layout is shared by the scene build and the kernels");
static_assert(csgray::kMaxGroupMembers <= static_cast<int32_t>(csgray::kWalkMaxChunks * csgray::kWalkChunk),
              "a group's chunk numbers fit the kernels' one-byte hit-chunk list");

extern "C" {

const char* csg_version(void) { return "csgray 0.1 (gfx950)"; }

const char* csg_last_error(void) { return g_lastError.c_str(); }

csg_status csg_scene_parse_host(const char* text, size_t len, csg_scene** out, char* err, size_t errlen) {
  if (!out) return Fail(CSG_E_ARG, "csg_scene_parse: out is NULL");
  *out = nullptr;
  if (!text && len) return Fail(CSG_E_ARG, "csg_scene_parse: text is NULL");
  csg_scene* s = new (std::nothrow) csg_scene();
  if (!s) return Fail(CSG_E_OOM, "csg_scene_parse: host allocation failed");
  try {
    s->tree = csgray::CSGTree::Parse(std::string(text ? text : "", len));
  } catch (const csgray::ParseError& e) {
    delete s;
    CopyErr(err, errlen, e.what());
    return Fail(CSG_E_PARSE, e.what());
  } catch (const std::bad_alloc&) {
    delete s;
    CopyErr(err, errlen, "out of host memory");
    return Fail(CSG_E_OOM, "csg_scene_parse: host allocation failed");
  }
  const int32_t nn = static_cast<int32_t>(s->tree.nodes.size());
  const int32_t np = static_cast<int32_t>(s->tree.primitives.size());
  std::string limit;
  csg_status code = CSG_OK;
  if (nn > csgray::kMaxNodes) {
    limit = "scene has " + std::to_string(nn) + " nodes; the limit is " + std::to_string(csgray::kMaxNodes);
    code = CSG_E_ARG;
  } else if (np > csgray::kMaxPrimitives) {
    limit = "scene has " + std::to_string(np) + " primitives; hit ids are int16, the limit is " +
            std::to_string(csgray::kMaxPrimitives);
    code = CSG_E_ARG;
  } else if (s->tree.opDepth > csgray::MaxSupportedDepth()) {
    limit = "operator depth " + std::to_string(s->tree.opDepth) + " exceeds the supported " +
            std::to_string(csgray::MaxSupportedDepth());
    code = CSG_E_DEPTH;
  }
  if (code != CSG_OK) {
    delete s;
    CopyErr(err, errlen, limit);
    return Fail(code, limit);
  }
  s->flat = csgray::Flatten(s->tree);
  *out = s;
  return CSG_OK;
}

csg_status csg_scene_upload(csg_scene* s) {
  if (!s) return Fail(CSG_E_ARG, "csg_scene_upload: NULL scene");
  if (s->dmem) return CSG_OK;
  const csg_status st = Upload(s);
  if (st != CSG_OK && s->dmem) {
    (void)hipFree(s->dmem);
    s->dmem = nullptr;
  }
  return st;
}

csg_status csg_scene_parse(const char* text, size_t len, csg_scene** out, char* err, size_t errlen) {
  csg_status st = csg_scene_parse_host(text, len, out, err, errlen);
  if (st != CSG_OK) return st;
  st = csg_scene_upload(*out);
  if (st != CSG_OK) {
    CopyErr(err, errlen, g_lastError);
    const std::string keep = g_lastError;
    delete *out;
    *out = nullptr;
    g_lastError = keep;
  }
  return st;
}

csg_status csg_scene_free(csg_scene* scene) {
  if (!scene) return CSG_OK;
  csg_status st = CSG_OK;
  if (scene->dmem) {
    const hipError_t e = hipFree(scene->dmem);
    if (e != hipSuccess) st = FailHip(e, "hipFree");
  }
  delete scene;
  return st;
}

csg_status csg_scene_info_get(const csg_scene* s, csg_scene_info* info) {
  if (!s || !info) return Fail(CSG_E_ARG, "csg_scene_info_get: NULL argument");
  std::memset(info, 0, sizeof(*info));
  info->nodes = s->flat.nNodes;
  info->primitives = s->flat.nPrims;
  info->op_depth = s->flat.opDepth;
  for (const auto& p : s->tree.primitives) {
    if (p.type == csgray::kSphere) info->spheres++;
    else if (p.type == csgray::kCylinder) info->cylinders++;
    else info->cubes++;
  }
  info->device_bytes = static_cast<int64_t>(s->bytes);
  info->prims_in_boxes = s->flat.conservative ? 1 : 0;
  info->union_groups = s->flat.unionGroups;
  return CSG_OK;
}

csg_status csg_scene_export(const csg_scene* s, int32_t* links, float* boxes, float* prims) {
  if (!s) return Fail(CSG_E_ARG, "csg_scene_export: NULL scene");
  const auto& nodes = s->tree.nodes;
  for (size_t i = 0; i < nodes.size(); i++) {
    if (links) {
      int32_t* l = links + 5 * i;
      l[0] = nodes[i].type; l[1] = nodes[i].primitiveIdx; l[2] = nodes[i].left;
      l[3] = nodes[i].right; l[4] = nodes[i].parent;
    }
    if (boxes) {
      float* b = boxes + 6 * i;
      for (int k = 0; k < 3; k++) {
        b[k] = nodes[i].box.lo[k];
        b[3 + k] = nodes[i].box.hi[k];
      }
    }
  }
  if (prims) {
    const auto& ps = s->tree.primitives;
    for (size_t i = 0; i < ps.size(); i++) {
      float* o = prims + 14 * i;
      const auto& p = ps[i];
      o[0] = static_cast<float>(p.id); o[1] = static_cast<float>(p.type);
      o[2] = p.x; o[3] = p.y; o[4] = p.z; o[5] = p.r; o[6] = p.g; o[7] = p.b;
      o[8] = p.radius; o[9] = p.height; o[10] = p.size;
      o[11] = p.axis[0]; o[12] = p.axis[1]; o[13] = p.axis[2];
    }
  }
  return CSG_OK;
}

void csg_camera_default(csg_camera* cam) {  // Camera() : Camera.h:18-20
  if (!cam) return;
  cam->pos[0] = 0; cam->pos[1] = 0; cam->pos[2] = 5;
  cam->pitch = 0; cam->yaw = 0;
  cam->vfov = 90.0f * 3.14159f / 180.0f;
}

csg_status csg_camera_look_at(const float eye[3], const float target[3], float vfov_deg, csg_camera* cam) {
  if (!eye || !target || !cam) return Fail(CSG_E_ARG, "csg_camera_look_at: NULL argument");
  // forward = (-sin yaw cos pitch, sin pitch, -cos yaw cos pitch) (Camera.cpp:10-12)
  const double fx = static_cast<double>(target[0]) - eye[0];
  const double fy = static_cast<double>(target[1]) - eye[1];
  const double fz = static_cast<double>(target[2]) - eye[2];
  const double len = std::sqrt(fx * fx + fy * fy + fz * fz);
  if (!(len > 0)) return Fail(CSG_E_ARG, "csg_camera_look_at: eye equals target");
  cam->pos[0] = eye[0]; cam->pos[1] = eye[1]; cam->pos[2] = eye[2];
  const float lim = csgray::PitchLimit();
  const float pitch = static_cast<float>(std::asin(fy / len));
  cam->pitch = std::fmax(-lim, std::fmin(lim, pitch));
  cam->yaw = static_cast<float>(std::atan2(-fx, -fz));
  cam->vfov = vfov_deg * 3.14159f / 180.0f;  // Camera::setFOV, Camera.h:49-52
  return CSG_OK;
}

csg_status csg_camera_basis(const csg_camera* cam, float out10[10]) {
  if (!cam || !out10) return Fail(CSG_E_ARG, "csg_camera_basis: NULL argument");
  const csgray::CameraBasis b = csgray::MakeBasis(cam->pos, cam->pitch, cam->yaw, cam->vfov);
  for (int k = 0; k < 3; k++) {
    out10[k] = b.forward[k];
    out10[3 + k] = b.right[k];
    out10[6 + k] = b.up[k];
  }
  out10[9] = b.tanHalf;
  return CSG_OK;
}

void csg_light_default(csg_light* light) {  // DirectionalLight.h:10-11
  if (!light) return;
  csgray::LightFromAngles(-60.f * 3.14159f / 180.f, -45.f * 3.14159f / 180.f, light->dir);
}

void csg_light_from_angles(float polar, float azimuth, csg_light* light) {
  if (!light) return;
  csgray::LightFromAngles(polar, azimuth, light->dir);
}

void csg_phong_default(csg_phong* ph) {  // RaycastingKernels.cu:71-74, 109
  if (!ph) return;
  ph->ka = 0.2f; ph->kd = 0.8f; ph->ks = 0.7f; ph->shininess = 30.0f;
  ph->bg[0] = 0.08f; ph->bg[1] = 0.08f; ph->bg[2] = 0.11f; ph->bg[3] = 1.0f;
}

}  // extern "C"

namespace csgray {
namespace rt {
csg_status ParseOpts(const csg_render_opts* opts, csgray::LaunchConfig* cfg) {
  if (!opts) return CSG_OK;
  if (opts->mode != CSG_MODE_EXACT && opts->mode != CSG_MODE_FAST) return Fail(CSG_E_ARG, "csg_render: unknown mode");
  cfg->mode = opts->mode == CSG_MODE_EXACT ? csgray::Mode::kExact : csgray::Mode::kFast;
  cfg->count = opts->counters != 0;
  if (cfg->count && !opts->counter_out) return Fail(CSG_E_ARG, "csg_render: counters requested without a buffer");
  if (opts->block != 0 && opts->block != 256 && opts->block != 512 && opts->block != 1024)
    return Fail(CSG_E_ARG, "csg_render: block must be 0, 256, 512 or 1024");
  if (opts->scene_placement < 0 || opts->scene_placement > 3)
    return Fail(CSG_E_ARG, "csg_render: scene_placement must be 0, 1, 2 or 3");
  if (opts->samples < 0 || opts->samples > 1024) return Fail(CSG_E_ARG, "csg_render: samples must be in [0, 1024]");
  if (opts->tile_walk != 0 && opts->tile_walk != 1) return Fail(CSG_E_ARG, "csg_render: tile_walk must be 0 or 1");
  cfg->block = opts->block;
  cfg->sceneInLds = opts->scene_placement == 0 ? -1
                    : opts->scene_placement == 2 ? 1
                    : opts->scene_placement == 3 ? 2 : 0;
  cfg->morton = opts->tile_walk == 1;
  cfg->budget = opts->coop_budget;
  if (opts->coop_when != 0 && opts->coop_when != 1) return Fail(CSG_E_ARG, "coop_when must be 0 or 1");
  cfg->coopEager = opts->coop_when == 1;
  if (opts->coop_grace > 100000) return Fail(CSG_E_ARG, "coop_grace above 100000 percent");
  cfg->coopGrace = opts->coop_grace;
  cfg->coopIntervals = opts->coop_intervals >= 0;
  return CSG_OK;
}

size_t SceneBytes(const csg_scene* s) {
  return (static_cast<size_t>(s->dev.nOps) * 4 + static_cast<size_t>(s->dev.nPrims) * 2) * 16;
}

// Validates one job and fills its device descriptor (outputs, camera basis,
// light, shading, tile counts).  `tileBase` is its first queue index.
csg_status BuildJob(const csg_scene* s, const csg_camera* cam, const csg_light* light, const csg_phong* phong,
                    int32_t width, int32_t height, const csg_outputs* out, const csg_render_opts* opts,
                    int dev, int64_t tileBase, csgray::Job* jb) {
  if (!s || !cam || !light || !phong || !out) return Fail(CSG_E_ARG, "csg_render: NULL argument");
  // The reference's ray generation divides by width-1 and height-1
  // (RaycastingKernels.cu:11-12), so both must be at least 2.
  if (width < 2 || height < 2) return Fail(CSG_E_ARG, "csg_render: width and height must be >= 2");
  const int32_t spp = opts && opts->samples > 1 ? opts->samples : 1;
  const int64_t rays = static_cast<int64_t>(width) * height * spp;
  if (rays > (int64_t{1} << 31)) return Fail(CSG_E_ARG, "csg_render: more than 2^31 rays");
  if (!s->dmem) return Fail(CSG_E_ARG, "csg_render: scene is not on a device (call csg_scene_upload)");
  if (dev != s->device) return Fail(CSG_E_ARG, "csg_render: scene was uploaded to another device");
  std::memset(jb, 0, sizeof(*jb));
  jb->sc = s->dev;
  // fast mode culls conservative scenes with their tight boxes
  const bool fast = !opts || opts->mode == CSG_MODE_FAST;
  if (fast && (!opts || opts->tight_boxes >= 0) && s->dev.ivOps) jb->sc.ops = s->dev.ivOps;
  const csgray::CameraBasis b = csgray::MakeBasis(cam->pos, cam->pitch, cam->yaw, cam->vfov);
  for (int k = 0; k < 3; k++) {
    jb->pos[k] = b.pos[k];
    jb->forward[k] = b.forward[k];
    jb->right[k] = b.right[k];
    jb->up[k] = b.up[k];
  }
  jb->tanHalf = b.tanHalf;
  csgray::NormalizeLight(light->dir, jb->light);
  jb->ka = phong->ka;
  jb->kd = phong->kd;
  jb->ks = phong->ks;
  jb->shininess = phong->shininess;
  for (int k = 0; k < 4; k++) jb->bg[k] = phong->bg[k];
  jb->width = width;
  jb->height = height;
  jb->tilesX = (width + 7) / 8;
  const int64_t tiles = static_cast<int64_t>(jb->tilesX) * ((height + 7) / 8) * spp;
  if (tileBase + tiles >= (int64_t{1} << 31)) return Fail(CSG_E_ARG, "csg_render: too many tiles in one launch");
  jb->tiles = static_cast<int32_t>(tiles);
  jb->tileBase = static_cast<int32_t>(tileBase);
  jb->spp = spp;
  jb->seed = opts ? opts->seed : 0u;
  jb->rgba = reinterpret_cast<float4*>(out->rgba);
  jb->depth = out->depth;
  jb->primId = out->prim_id;
  jb->flags = out->flags;
  jb->counters = (opts && opts->counters) ? opts->counter_out : nullptr;
  return CSG_OK;
}

// Per-(device, stream) job-table buffers for batched launches: a device
// table and a pinned host staging copy, grown on demand.  `copied` marks the
// last upload; the staging copy is rewritten only after it completes.
struct JobTable {
  void* dev = nullptr;
  void* host = nullptr;
  size_t cap = 0;
  hipEvent_t copied = nullptr;
};
std::mutex g_jobsMutex;
std::map<std::pair<int, hipStream_t>, JobTable> g_jobTables;

csg_status UploadJobs(int dev, hipStream_t stream, const csgray::Job* jobs, size_t n, const csgray::Job** out) {
  std::lock_guard<std::mutex> lock(g_jobsMutex);
  JobTable& t = g_jobTables[{dev, stream}];
  const size_t bytes = sizeof(csgray::Job) * n;
  hipError_t e;
  if (t.cap < bytes) {
    if (t.dev) {
      // earlier launches on this stream may still read the old table
      e = hipStreamSynchronize(stream);
      if (e != hipSuccess) return FailHip(e, "hipStreamSynchronize");
      (void)hipFree(t.dev);
      (void)hipHostFree(t.host);
      t.dev = t.host = nullptr;
      t.cap = 0;
    }
    const size_t cap = std::max<size_t>(bytes, 64 * 1024);
    e = hipMalloc(&t.dev, cap);
    if (e != hipSuccess) return e == hipErrorOutOfMemory ? Fail(CSG_E_OOM, "job table allocation failed")
                                                         : FailHip(e, "hipMalloc");
    e = hipHostMalloc(&t.host, cap, hipHostMallocDefault);
    if (e != hipSuccess) {
      (void)hipFree(t.dev);
      t.dev = nullptr;
      return Fail(CSG_E_OOM, "pinned job staging allocation failed");
    }
    t.cap = cap;
    if (!t.copied) {
      e = hipEventCreateWithFlags(&t.copied, hipEventDisableTiming);
      if (e != hipSuccess) return FailHip(e, "hipEventCreate");
    }
  } else if (t.copied) {
    e = hipEventSynchronize(t.copied);  // the staging copy is free once the last upload is done
    if (e != hipSuccess) return FailHip(e, "hipEventSynchronize");
  }
  std::memcpy(t.host, jobs, bytes);
  e = hipMemcpyAsync(t.dev, t.host, bytes, hipMemcpyHostToDevice, stream);
  if (e != hipSuccess) return FailHip(e, "hipMemcpyAsync of the job table");
  e = hipEventRecord(t.copied, stream);
  if (e != hipSuccess) return FailHip(e, "hipEventRecord");
  *out = static_cast<const csgray::Job*>(t.dev);
  return CSG_OK;
}

size_t StreamWorkspaceBytes(int32_t maxJobs) {
  const size_t table = std::max<size_t>(sizeof(csgray::Job) * static_cast<size_t>(std::max(maxJobs, 1)), 64 * 1024);
  return Align256(csgray::kQueueBytes) + Align256(csgray::kHeavyBytes) + Align256(table);
}

csg_status ReleaseStreamWorkspace(int dev, hipStream_t stream) {
  csg_status st = CSG_OK;
  const std::pair<int, hipStream_t> key{dev, stream};
  {
    std::lock_guard<std::mutex> lock(g_queueMutex);
    auto it = g_queues.find(key);
    if (it != g_queues.end()) {
      if (hipFree(it->second) != hipSuccess) st = Fail(CSG_E_HIP, "hipFree of a queue failed");
      g_queues.erase(it);
    }
  }
  {
    std::lock_guard<std::mutex> lock(g_heavyMutex);
    auto it = g_heavy.find(key);
    if (it != g_heavy.end()) {
      if (hipFree(it->second) != hipSuccess) st = Fail(CSG_E_HIP, "hipFree of a heavy-ray list failed");
      g_heavy.erase(it);
    }
  }
  {
    std::lock_guard<std::mutex> lock(g_jobsMutex);
    auto it = g_jobTables.find(key);
    if (it != g_jobTables.end()) {
      JobTable& t = it->second;
      if (t.copied) (void)hipEventSynchronize(t.copied);
      if (t.dev && hipFree(t.dev) != hipSuccess) st = Fail(CSG_E_HIP, "hipFree of a job table failed");
      if (t.host) (void)hipHostFree(t.host);
      if (t.copied) (void)hipEventDestroy(t.copied);
      g_jobTables.erase(it);
    }
  }
  return st;
}
}  // namespace rt
}  // namespace csgray

extern "C" {

csg_status csg_render_plan(const csg_scene* s, const csg_render_opts* opts, csg_launch_info* info) {
  if (!s || !info) return Fail(CSG_E_ARG, "csg_render_plan: NULL argument");
  csgray::LaunchConfig cfg;
  const csg_status st = ParseOpts(opts, &cfg);
  if (st != CSG_OK) return st;
  int cus = 0;
  size_t ldsLimit = 0;
  hipError_t e = csgray::DeviceShapeLimits(&cus, &ldsLimit);
  if (e != hipSuccess) return FailHip(e, "device query");
  csgray::LaunchShape shape;
  e = csgray::ChooseShape(s->dev.opDepth, SceneBytes(s), cfg, &shape, 0, 0, s->dev.nOps);
  if (e != hipSuccess) return FailHip(e, "choosing the launch shape");
  info->block = shape.block;
  info->grid = shape.grid;
  info->lds_bytes = static_cast<int64_t>(shape.lds);
  info->scene_in_lds = shape.sceneInLds ? 1 : 0;
  info->cache_ops = shape.cacheOps;
  info->compute_units = cus;
  info->lds_limit = static_cast<int64_t>(ldsLimit);
  const bool lists = CoopLists(cfg, s->dev.conservative != 0);
  const CoopPlan c = PlanCoop(cfg, s->dev.nOps, s->dev.opLeaves, lists);
  info->coop_budget = c.budget;
  info->coop_grid = static_cast<int32_t>(c.waves);
  info->coop_block = 64;
  // the grace of a 1920x1080 frame (240 x 135 tiles)
  const int64_t hdTiles = 240 * 135;
  const int64_t waves = int64_t{shape.grid} * (shape.block / 64);
  info->coop_grace = c.budget > 0 ? (cfg.coopGrace > 0 ? cfg.coopGrace
                                                       : (cfg.coopGrace < 0 ? 0
                                                                            : (DefaultCoopGraceMille(waves, hdTiles, false,
                                                                                                     lists) + 5) / 10))
                                  : 0;
  info->coop_lds_bytes = static_cast<int64_t>(c.ldsBytes);
  return CSG_OK;
}

namespace {
// Reads a 64-bit running total of the stream's queue words after waiting for the stream.
csg_status QueueTotal(void* stream, uint32_t word, int64_t* value) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return FailHip(e, "hipGetDevice");
  const hipStream_t hs = static_cast<hipStream_t>(stream);
  uint32_t* q = nullptr;
  const csg_status qs = QueueFor(dev, hs, &q);
  if (qs != CSG_OK) return qs;
  e = hipStreamSynchronize(hs);
  if (e != hipSuccess) return FailHip(e, "hipStreamSynchronize");
  unsigned long long v = 0;
  e = hipMemcpy(&v, q + word, sizeof v, hipMemcpyDeviceToHost);
  if (e != hipSuccess) return FailHip(e, "hipMemcpy");
  *value = static_cast<int64_t>(v);
  return CSG_OK;
}
}  // namespace

csg_status csg_coop_interval_rays(void* stream, int64_t* rays) {
  if (!rays) return Fail(CSG_E_ARG, "csg_coop_interval_rays: NULL argument");
  return QueueTotal(stream, 8, rays);
}

csg_status csg_coop_rays(void* stream, int64_t* rays) {
  if (!rays) return Fail(CSG_E_ARG, "csg_coop_rays: NULL argument");
  return QueueTotal(stream, 6, rays);
}

float csg_jitter(uint32_t seed, uint32_t pixel, uint32_t sample, uint32_t axis) {
  return csgray::JitterUnit(csgray::JitterHash(seed, pixel, sample, axis));
}

csg_status csg_render_ex(const csg_scene* s, const csg_camera* cam, const csg_light* light, const csg_phong* phong,
                         int32_t width, int32_t height, const csg_outputs* out, const csg_render_opts* opts,
                         void* stream) {
  csgray::LaunchConfig cfg;
  const csg_status ost = ParseOpts(opts, &cfg);
  if (ost != CSG_OK) return ost;
  if (!s || !cam || !light || !phong || !out) return Fail(CSG_E_ARG, "csg_render: NULL argument");
  if (!s->dmem) return Fail(CSG_E_ARG, "csg_render: scene is not on a device (call csg_scene_upload)");
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
  const csg_status qs = QueueFor(dev, hs, &p.queue);
  if (qs != CSG_OK) return qs;
  uint2* heavy = nullptr;
  if (cfg.budget >= 0) {
    const csg_status hsd = HeavyFor(dev, hs, &heavy);
    if (hsd != CSG_OK) return hsd;
  }
  PrepareCoop(cfg, s->dev.nOps, s->dev.opLeaves, s->dev.conservative != 0, heavy, shape, &p);
  e = csgray::LaunchRender(p, cfg, shape, hs);
  if (e != hipSuccess) return FailHip(e, "render launch");
  return CSG_OK;
}

csg_status csg_render(const csg_scene* s, const csg_camera* cam, const csg_light* light, const csg_phong* phong,
                      int32_t width, int32_t height, const csg_outputs* out, void* stream) {
  return csg_render_ex(s, cam, light, phong, width, height, out, nullptr, stream);
}

csg_status csg_batch_order(const csg_job* jobs, int32_t njobs, int32_t* order) {
  if (!jobs || !order || njobs < 0) return Fail(CSG_E_ARG, "csg_batch_order: bad argument");
  std::vector<std::pair<double, int32_t>> cost(static_cast<size_t>(njobs));
  for (int32_t i = 0; i < njobs; i++) {
    const double depth = jobs[i].scene ? jobs[i].scene->dev.opDepth : 0;
    cost[i] = {-static_cast<double>(jobs[i].width) * jobs[i].height * (1.0 + depth), i};
  }
  std::stable_sort(cost.begin(), cost.end());
  for (int32_t i = 0; i < njobs; i++) order[i] = cost[i].second;
  return CSG_OK;
}

csg_status csg_render_batch(const csg_job* jobs, int32_t njobs, const csg_render_opts* opts, void* stream) {
  if (!jobs || njobs <= 0) return Fail(CSG_E_ARG, "csg_render_batch: no jobs");
  if (opts && opts->tile_order) return Fail(CSG_E_ARG, "csg_render_batch: tile_order applies to single renders");
  csgray::LaunchConfig cfg;
  const csg_status ost = ParseOpts(opts, &cfg);
  if (ost != CSG_OK) return ost;
  if (cfg.sceneInLds >= 1) return Fail(CSG_E_ARG, "csg_render_batch: scenes of a batch are read from global memory");
  cfg.sceneInLds = 0;
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return FailHip(e, "hipGetDevice");
  const hipStream_t hs = static_cast<hipStream_t>(stream);
  csgray::RenderParams p{};
  int32_t depth = 1;
  int64_t base = 0;
  for (int32_t start = 0; start < njobs; start += csgray::kMaxBatchJobs) {
    const int32_t n = std::min(njobs - start, csgray::kMaxBatchJobs);
    std::vector<csgray::Job> host(static_cast<size_t>(n));
    base = 0;
    depth = 1;
    for (int32_t i = 0; i < n; i++) {
      const csg_job& j = jobs[start + i];
      const csg_status js = BuildJob(j.scene, &j.camera, &j.light, &j.phong, j.width, j.height, &j.out, opts, dev,
                                     base, &host[i]);
      if (js != CSG_OK) return js;
      base += host[i].tiles;
      depth = std::max(depth, j.scene->dev.opDepth);
    }
    const csg_status bs = UploadJobs(dev, hs, host.data(), host.size(), &p.jobs);
    if (bs != CSG_OK) return bs;
    p.numJobs = n;
    p.numTiles = static_cast<int32_t>(base);
    p.stackSlots = depth;
    p.morton = cfg.morton ? 1 : 0;
    csgray::LaunchShape shape;
    e = csgray::ChooseShape(depth, 0, cfg, &shape, n);
    if (e != hipSuccess) return FailHip(e, "choosing the launch shape");
    csgray::FitGridToTiles(&shape, p.numTiles);
    p.totalWaves = static_cast<uint32_t>(shape.grid) * static_cast<uint32_t>(shape.block / 64);
    const csg_status qs = QueueFor(dev, hs, &p.queue);
    if (qs != CSG_OK) return qs;
    uint2* heavy = nullptr;
    if (cfg.budget >= 0) {
      const csg_status hsd = HeavyFor(dev, hs, &heavy);
      if (hsd != CSG_OK) return hsd;
    }
    int32_t maxOps = 1, maxLeaves = 1;
    bool anyInBoxes = false;
    for (int32_t i = 0; i < n; i++) {
      maxOps = std::max(maxOps, jobs[start + i].scene->dev.nOps);
      maxLeaves = std::max(maxLeaves, jobs[start + i].scene->dev.opLeaves);
      anyInBoxes |= jobs[start + i].scene->dev.conservative != 0;
    }
    PrepareCoop(cfg, maxOps, maxLeaves, anyInBoxes, heavy, shape, &p);
    e = csgray::LaunchRender(p, cfg, shape, hs);
    if (e != hipSuccess) return FailHip(e, "batched render launch");
  }
  return CSG_OK;
}

csg_status csg_sync(void* stream) {
  const hipError_t e = hipStreamSynchronize(static_cast<hipStream_t>(stream));
  if (e != hipSuccess) return FailHip(e, "hipStreamSynchronize");
  return CSG_OK;
}

}  // extern "C"
