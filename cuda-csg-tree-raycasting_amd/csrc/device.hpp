// Device-side scene view and launch parameters shared by the kernels and the
// host runtime.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace csgray {

// Read-only scene arrays in device memory (layout in scene.hpp: FlatScene).
struct DevScene {
  const float4* ops;    // 4 per operator
  const float4* primA;  // centre, radius / edge
  const float4* primB;  // axis, height
  const float4* primC;  // colour
  uint32_t rootDesc;    // descriptor of the root node
  int32_t nOps;
  int32_t nPrims;
  int32_t opDepth;
  int32_t opLeaves;  // operators without an operator child: a ray's live cooperative tasks never exceed it
  uint32_t conservative;  // 1: every primitive lies inside its box (interval classification applies)
  const float4* ivOps;    // operator records with fast mode's culling boxes (tight where the reference allows)
  const float4* walkOps;  // interval classification's records: ivOps with Union groups flattened (scene.hpp)
};

// One render job: a scene seen by one camera at one resolution, with its
// output buffers.  A plain render passes one job by value as the kernel
// argument; a batched render reads a device array of them.
struct Job {
  DevScene sc;
  float pos[3];
  float forward[3], right[3], up[3];
  float tanHalf;
  float light[3];  // normalised, toward the light
  float ka, kd, ks, shininess;
  float bg[4];
  int32_t width, height;
  int32_t tilesX;     // 8x8-pixel tiles per row
  int32_t tiles;      // tiles of this job (all samples)
  int32_t tileBase;   // first queue index of this job (batched)
  int32_t spp;        // rays per pixel; sample s of pixel p is output entry p*spp + s
  uint32_t seed;      // jitter seed (spp > 1)
  int32_t pad_;
  float4* rgba;
  float* depth;
  int32_t* primId;
  uint8_t* flags;
  int32_t* counters;  // 6 per ray when counting
};

// Launch-wide parameters.
struct RenderParams {
  Job job;                    // plain: the one job
  const Job* jobs;            // batched: device array, ascending tileBase
  int32_t numJobs;            // 0 = plain
  int32_t numTiles;           // queue length (all jobs)
  int32_t stackSlots;         // LDS stack slots per lane (max operator depth of the launch)
  uint32_t totalWaves;        // waves in the grid (sets the default tail grace)
  uint32_t* queue;            // work-queue words (layout below); zero between launches
  const uint32_t* tileOrder;  // optional: order in which queue entries map to tiles (plain)
  int32_t morton;             // 1: tiles of a job are visited in Morton order
  // Cooperative evaluation of heavy rays: a ray still running after `budget`
  // steps once every tile has been taken (at once when `coopEager`) is
  // appended to `heavy` ({job, output index}); a second kernel evaluates each
  // one with a whole wave.  budget <= 0: never defer.
  int32_t budget;
  int32_t coopEager;
  int32_t coopGrace;          // tail mode: per mille of the tile hand-out time to wait after the last tile
  uint2* heavy;
  uint32_t heavyCap;
  int32_t coopOps;            // operator-state slots per wave (max operators of the launch)
  int32_t coopLists;          // 1: the region holds interval classification's view too (fast mode)
  uint32_t coopQueueCap;      // task ring entries (power of two >= the launch's largest opLeaves)
  uint32_t coopWaves;         // waves in the cooperative grid
  // Resume area (`resumeBytes`, right after the heavy list): for the first
  // ResumeCap() handed-over rays, the state machine's state at hand-over, so
  // the cooperative kernel continues the ray instead of restarting it; its
  // last eighth the first DirCap() handed-over rays' directions.
  uint32_t* resume;
  uint32_t resumeBytes;
  int32_t cacheOps;           // plain global-memory launches: operator records [0, cacheOps) staged in LDS
};
// queue words: [1] render workgroups done, [2] heavy rays, [4] coop waves done, [6..7] heavy rays handed
// over on this stream so far (64-bit), [8..9] of those, the ones finished from boundary lists; on their own cache line, [32..33] time the first tile was taken,
// [34..35] time the tiles ran out (64-bit wall clock | 1, 0 = not yet).  Tiles and heavy rays are taken
// from kQueueStripes counters each, 256 B apart: stripe s hands out items s, s + kQueueStripes, ...
// (counter = items of the stripe taken so far), so the waves' atomics spread over many addresses
// instead of queueing on one.
constexpr uint32_t kQueueStripes = 32;
constexpr uint32_t kStripeWords = 64;
constexpr uint32_t kTileStripeWord = 64;
constexpr uint32_t kHeavyStripeWord = kTileStripeWord + kQueueStripes * kStripeWords;
constexpr size_t kQueueBytes = sizeof(uint32_t) * (kHeavyStripeWord + kQueueStripes * kStripeWords);

// Jitter of sample s of pixel p: a 32-bit integer hash of (seed, p, s)
// mapped to [0, 1) with 24 bits; identical on host and device.
__host__ __device__ inline uint32_t JitterHash(uint32_t seed, uint32_t pixel, uint32_t sample, uint32_t axis) {
  uint32_t h = seed ^ (pixel * 0x9E3779B1u) ^ (sample * 0x85EBCA77u) ^ (axis * 0xC2B2AE3Du);
  h ^= h >> 16;
  h *= 0x7FEB352Du;
  h ^= h >> 15;
  h *= 0x846CA68Bu;
  h ^= h >> 16;
  return h;
}
__host__ __device__ inline float JitterUnit(uint32_t h) { return static_cast<float>(h >> 8) * (1.0f / 16777216.0f); }

enum class Mode : int32_t { kExact = 0, kFast = 1 };

struct LaunchConfig {
  Mode mode = Mode::kFast;
  bool count = false;
  int32_t block = 0;   // 0 = choose
  int32_t sceneInLds = -1;  // -1 = choose, 0 = global, 1 = LDS, 2 = global with the top of the tree cached in LDS
  bool morton = false;      // visit each job's tiles in Morton order
  int32_t budget = 0;       // 0: default step budget, < 0: no deferral, > 0: steps before deferral
  bool coopEager = false;   // hand rays over as soon as the budget is spent, not only after the queue drains
  int32_t coopGrace = 0;    // 0: default; < 0: none; else percent (see RenderParams::coopGrace)
  bool coopIntervals = true;  // interval classification of heavy rays where it applies
};

// Launch geometry chosen for a scene: workgroup size, grid, LDS bytes.
struct LaunchShape {
  int32_t block = 256;
  int32_t grid = 0;
  size_t lds = 0;
  bool sceneInLds = false;
  int32_t cacheOps = 0;  // operator records cached in LDS (scene in global memory)
};

// Fewer workgroups when the launch has fewer tiles than the grid has waves:
// the extra waves would only pop an empty queue (and stage the scene).
inline void FitGridToTiles(LaunchShape* s, int64_t tiles) {
  const int64_t waves = s->block / 64;
  const int64_t need = (tiles + waves - 1) / waves;
  if (need < s->grid) s->grid = static_cast<int>(need > 1 ? need : 1);
}

// Picks the launch shape on the current device for a launch whose scenes
// have at most `depth` operator levels; `sceneBytes` > 0 allows staging that
// many bytes of scene in LDS (single-scene launches only).  `tiles` is the
// launch's tile count and `ops` its scene's operator count (0: unknown,
// taken as a large frame and a large tree).
hipError_t ChooseShape(int32_t depth, size_t sceneBytes, const LaunchConfig& cfg, LaunchShape* shape,
                       int32_t batchJobs = 0, int64_t tiles = 0, int32_t ops = 0);

// Compute units, LDS bytes per workgroup and per CU of the current device.
hipError_t DeviceShapeLimits(int* computeUnits, size_t* ldsPerBlock, size_t* ldsPerCu = nullptr);

// Enqueues one launch on `stream` with the given shape (p.numJobs == 0:
// plain render of p.job; otherwise the batch in p.jobs).
hipError_t LaunchRender(const RenderParams& p, const LaunchConfig& cfg, const LaunchShape& shape,
                        hipStream_t stream);

// Most jobs one batched launch takes (their tile prefix is staged in LDS).
constexpr int32_t kMaxBatchJobs = 1024;

// Default per-lane step budget before a ray is handed to the cooperative
// kernel (with the state machine there, and where it classifies rays by
// intervals: cheese frames of 1280x720 to 3840x2160 measured best or level
// at 64 steps against 160), and the capacity of the heavy-ray list.
constexpr int32_t kDefaultBudget = 384;
constexpr int32_t kDefaultIntervalBudget = 64;
constexpr uint32_t kHeavyCap = 1u << 22;
// Resume records of handed-over rays: 20 words each (current operator,
// action, stack slots, tmin, the L/R/held records, the event counts so far),
// then per ray and stack slot the slot's entry (8 B) and the operator it
// belongs to (4 B).  kResumeFlag in a heavy-list entry's job word marks a
// ray with a record.  The last eighth of the area holds the directions of
// the first handed-over rays, 16 B a slot (kDirFlag marks a ray with one),
// whether or not they left their state: the cooperative kernel starts from
// it instead of building the ray again, and neighbouring slots share cache
// lines.  The records take the rest.
constexpr size_t kResumeBytes = size_t{64} << 20;
constexpr uint32_t kResumeWords = 20;
constexpr uint32_t kResumeFlag = 1u << 31;
constexpr uint32_t kDirFlag = 1u << 30;
constexpr uint32_t kJobWordIndex = ~(kResumeFlag | kDirFlag);
constexpr size_t ResumeDirBytes(size_t resumeBytes) { return (resumeBytes / 8) & ~size_t{15}; }
// Bytes of one heavy list together with its resume area.
constexpr size_t kHeavyBytes = sizeof(uint32_t) * 2 * kHeavyCap + kResumeBytes;

// LDS bytes the cooperative kernel needs for scenes of at most `ops`
// operators of which at most `leaves` have no operator child, with room for
// interval classification when `lists` (0 when it cannot run: the caller
// then disables deferral).
size_t CoopLdsBytes(int32_t ops, int32_t leaves, bool lists, uint32_t* queueCap);
// LDS of one wave's cooperative state, i.e. of one cooperative workgroup,
// rounded to 16 B.  Two views of the same bytes: the state machine's (one
// 16-byte state per operator, then the task ring's float tmins) and, when
// `lists`, interval classification's (per operator two 4-byte child masks,
// a 4-byte word and a 2-byte entry slot; then the crossings, 12 bytes each,
// the pending primitive tests, 8 bytes each, and two rounds' hit chunks);
// the ring's 16-bit tasks end the region in both.
#if defined(__HIPCC__)
#define CSG_HOST_DEVICE __host__ __device__
#else
#define CSG_HOST_DEVICE
#endif
constexpr size_t kCoopOpStateBytes = 16;
constexpr size_t kCoopTaskBytes = 6;
constexpr uint32_t kCoopMaxCross = 31;  // crossings a ray's interval masks (32 bits) can order
constexpr size_t kCoopJobs = 192;       // pending primitive tests: fewer than 64 left over + two a lane
// Union groups in the interval walk's records (scene.hpp): the type-word flag
// of a group's record, the members one chunk box covers, and the hit chunks
// of two rounds of chunk tests (kept in LDS, one byte each: a group has at
// most kWalkMaxChunks chunks).
constexpr uint32_t kWalkGroupFlag = 0x200u;
constexpr uint32_t kWalkChunk = 4;
constexpr uint32_t kWalkMaxChunks = 256;
constexpr size_t kCoopHitChunks = 128;
// Rays the cooperative kernel has finished wait after the region (heavy slot,
// hit time, hit bits; 16 B each) until kCoopShadeBatch of them are shaded at
// once, one a lane; then 32 B of state for a ray classified in stretches.
constexpr uint32_t kCoopShadeBatch = 14;
constexpr size_t kCoopDoneBytes = 16 * (kCoopShadeBatch + 2);
// Interval classification's crossings start after the per-operator state
// (14 B an operator), 8-byte aligned for the primitive tests that follow them.
CSG_HOST_DEVICE inline size_t CoopCrossOffset(int32_t ops) {
  return (static_cast<size_t>(ops) * 14 + 7) & ~static_cast<size_t>(7);
}
CSG_HOST_DEVICE inline size_t CoopWaveLds(int32_t ops, uint32_t queueCap, int32_t lists) {
  const size_t machine =
      static_cast<size_t>(ops) * kCoopOpStateBytes + static_cast<size_t>(queueCap) * kCoopTaskBytes;
  const size_t masks = CoopCrossOffset(ops) + 12 * (static_cast<size_t>(kCoopMaxCross) + 1) + 8 * kCoopJobs +
                       kCoopHitChunks + static_cast<size_t>(queueCap) * 2;
  const size_t bytes = lists && masks > machine ? masks : machine;
  return (bytes + 15) & ~static_cast<size_t>(15);
}
size_t CoopLaunchLds(const RenderParams& p);

// Largest operator depth the traversal stack supports.
int32_t MaxSupportedDepth();

}  // namespace csgray
