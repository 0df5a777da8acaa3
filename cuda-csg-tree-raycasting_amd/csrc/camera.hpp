// Camera and directional-light models of the reference, evaluated on the host
// once per render.  The kernels receive the finished basis, tan(fov/2) and the
// normalised light direction as arguments.
//
// Camera:            RenderManager/Camera/Camera.h:3-54, Camera.cpp:4-37
// DirectionalLight:  RenderManager/DirectionalLight.h:8-18
#pragma once

#include <cmath>

namespace csgray {

struct CameraBasis {
  float pos[3];
  float forward[3], right[3], up[3];
  float tanHalf;  // tan(fov / 2), evaluated once (the kernel evaluates it per pixel; same value)
};

// Camera::normalizeVector (Camera.cpp:29-37): divide by the length when it is
// positive, leave the vector alone otherwise.
inline void NormalizeInPlace(float* v) {
  const float len = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
  if (len > 0) {
    v[0] /= len;
    v[1] /= len;
    v[2] /= len;
  }
}

// Pitch limit of Camera::setRotation (Camera.h:28-35), in float as written.
inline float PitchLimit() { return 89.0f * 3.14159f / 180.0f; }

// Camera::updateVectors (Camera.cpp:4-27) for a camera at `pos` with pitch
// (rotX) clamped to +-89 degrees, yaw (rotY) and vertical fov in radians.
inline CameraBasis MakeBasis(const float pos[3], float pitch, float yaw, float vfov) {
  CameraBasis c;
  c.pos[0] = pos[0];
  c.pos[1] = pos[1];
  c.pos[2] = pos[2];
  const float lim = PitchLimit();
  const float rx = std::fmax(-lim, std::fmin(lim, pitch));
  const float ry = yaw;
  c.forward[0] = -std::sin(ry) * std::cos(rx);
  c.forward[1] = std::sin(rx);
  c.forward[2] = -std::cos(ry) * std::cos(rx);
  NormalizeInPlace(c.forward);
  c.right[0] = std::cos(ry);
  c.right[1] = 0;
  c.right[2] = -std::sin(ry);
  NormalizeInPlace(c.right);
  // up = forward x right, so (0, -1, 0) at pitch = yaw = 0: rays of image row 0
  // point down, and row 0 is the bottom scanline once displayed.
  c.up[0] = c.forward[1] * c.right[2] - c.forward[2] * c.right[1];
  c.up[1] = c.forward[2] * c.right[0] - c.forward[0] * c.right[2];
  c.up[2] = c.forward[0] * c.right[1] - c.forward[1] * c.right[0];
  NormalizeInPlace(c.up);
  c.tanHalf = std::tan(vfov / 2.0f);
  return c;
}

// Float3Utils.cuh:33-37 normalize(): multiply by 1/sqrt(a.a).  The shading
// kernel normalises the light direction per pixel (RaycastingKernels.cu:78);
// the value is identical for every pixel, so it is done once here.
inline void NormalizeLight(const float in[3], float out[3]) {
  const float invLen = 1.0f / std::sqrt(in[0] * in[0] + in[1] * in[1] + in[2] * in[2]);
  out[0] = invLen * in[0];
  out[1] = invLen * in[1];
  out[2] = invLen * in[2];
}

// DirectionalLight::getLightDir (DirectionalLight.h:13-18).
inline void LightFromAngles(float polar, float azimuth, float out[3]) {
  out[0] = std::sin(polar) * std::cos(azimuth);
  out[1] = std::cos(polar);
  out[2] = std::sin(polar) * std::sin(azimuth);
}

}  // namespace csgray
