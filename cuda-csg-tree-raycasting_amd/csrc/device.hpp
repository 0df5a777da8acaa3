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
};

// Everything one render needs, passed by value as the kernel argument.
struct RenderParams {
  DevScene sc;
  float pos[3];
  float forward[3], right[3], up[3];
  float tanHalf;
  float light[3];  // normalised, toward the light
  float ka, kd, ks, shininess;
  float bg[4];
  int32_t width, height;
  int32_t tilesX, numTiles;  // 8x8-pixel tiles, row-major
  uint32_t totalWaves;       // waves in the grid (queue reset bookkeeping)
  uint32_t* queue;           // [0] next tile, [1] waves finished; zero between launches
  const uint32_t* tileOrder; // optional: the order in which tiles are taken (numTiles entries)
  float4* rgba;
  float* depth;
  int32_t* primId;
  uint8_t* flags;
  int32_t* counters;  // 6 per pixel when counting
};

enum class Mode : int32_t { kExact = 0, kFast = 1 };

struct LaunchConfig {
  Mode mode = Mode::kFast;
  bool count = false;
  int32_t block = 0;   // 0 = choose
  int32_t sceneInLds = -1;  // -1 = choose, 0 = global, 1 = LDS
};

// Launch geometry chosen for a scene: workgroup size, grid, LDS bytes.
struct LaunchShape {
  int32_t block = 256;
  int32_t grid = 0;
  size_t lds = 0;
  bool sceneInLds = false;
};

// Picks the launch shape for `sc` on the current device.
hipError_t ChooseShape(const DevScene& sc, const LaunchConfig& cfg, LaunchShape* shape);

// Compute units and LDS bytes per workgroup of the current device.
hipError_t DeviceShapeLimits(int* computeUnits, size_t* ldsPerBlock);

// Enqueues one render on `stream` with the given shape.
hipError_t LaunchRender(const RenderParams& p, const LaunchConfig& cfg, const LaunchShape& shape,
                        hipStream_t stream);

// Largest operator depth the traversal stack supports.
int32_t MaxSupportedDepth();

}  // namespace csgray
