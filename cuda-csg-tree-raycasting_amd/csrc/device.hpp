// Device-side scene view and launch parameters shared by the kernels and the
// host runtime.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace csgray {

// Read-only scene arrays in device memory (layout in scene.hpp: FlatScene).
struct DevScene {
  const float4* nodeLo;
  const float4* nodeHi;
  const float4* primA;
  const float4* primB;
  const float4* primC;
  int32_t nNodes;
  int32_t nPrims;
  int32_t opDepth;
  int32_t pad_;
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
  float4* rgba;
  float* depth;
  int32_t* primId;
  uint8_t* flags;
  int32_t* counters;  // 6 per pixel when counting
};

enum class Mode : int32_t { kExact = 0, kFast = 1 };

struct LaunchConfig {
  Mode mode = Mode::kFast;
  int32_t block = 256;
  bool count = false;
};

// Enqueues one render on `stream`; returns the launch error code.
hipError_t LaunchRender(const RenderParams& p, const LaunchConfig& cfg, hipStream_t stream);

// Largest operator depth the traversal stack of one workgroup holds.
int32_t MaxSupportedDepth();

}  // namespace csgray
