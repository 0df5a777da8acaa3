// Headless output (PPM, PNG, raw arrays) and the interactive application's
// camera control, for scripted renders.  Host code only.
//
// Camera::move / rotate / setRotation: RenderManager/Camera/Camera.h:28-47
// Application::Input:                   Application.cpp:85-119
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <vector>

#include "camera.hpp"
#include "csg.h"
#include "runtime.hpp"

using csgray::rt::Fail;

namespace {

uint8_t ToByte(float v) {
  if (!(v > 0.0f)) return 0;  // also NaN
  if (v >= 1.0f) return 255;
  return static_cast<uint8_t>(v * 255.0f + 0.5f);
}

struct FileCloser {
  void operator()(FILE* f) const {
    if (f) std::fclose(f);
  }
};
using File = std::unique_ptr<FILE, FileCloser>;

csg_status OpenForWrite(const char* path, File* out) {
  if (!path) return Fail(CSG_E_ARG, "output path is NULL");
  out->reset(std::fopen(path, "wb"));
  if (!*out) return Fail(CSG_E_ARG, std::string("cannot open ") + path + " for writing");
  return CSG_OK;
}

csg_status CheckImage(const void* data, int32_t w, int32_t h) {
  if (!data) return Fail(CSG_E_ARG, "image data is NULL");
  if (w <= 0 || h <= 0 || static_cast<int64_t>(w) * h > (int64_t{1} << 31))
    return Fail(CSG_E_ARG, "image size out of range");
  return CSG_OK;
}

csg_status WriteAll(FILE* f, const void* p, size_t n) {
  if (n && std::fwrite(p, 1, n, f) != n) return Fail(CSG_E_ARG, "short write");
  return CSG_OK;
}

// CRC-32 (ISO 3309, reflected polynomial 0xEDB88320) as PNG chunks use it.
uint32_t Crc32(uint32_t crc, const uint8_t* p, size_t n) {
  static uint32_t table[256];
  static bool init = false;
  if (!init) {
    for (uint32_t i = 0; i < 256; i++) {
      uint32_t c = i;
      for (int k = 0; k < 8; k++) c = (c & 1) ? 0xEDB88320u ^ (c >> 1) : c >> 1;
      table[i] = c;
    }
    init = true;
  }
  crc = ~crc;
  for (size_t i = 0; i < n; i++) crc = table[(crc ^ p[i]) & 0xFF] ^ (crc >> 8);
  return ~crc;
}

void PutBe32(std::vector<uint8_t>& v, uint32_t x) {
  v.push_back(static_cast<uint8_t>(x >> 24));
  v.push_back(static_cast<uint8_t>(x >> 16));
  v.push_back(static_cast<uint8_t>(x >> 8));
  v.push_back(static_cast<uint8_t>(x));
}

csg_status WriteChunk(FILE* f, const char type[4], const std::vector<uint8_t>& data) {
  std::vector<uint8_t> buf;
  buf.reserve(data.size() + 12);
  PutBe32(buf, static_cast<uint32_t>(data.size()));
  buf.insert(buf.end(), type, type + 4);
  buf.insert(buf.end(), data.begin(), data.end());
  PutBe32(buf, Crc32(0, buf.data() + 4, data.size() + 4));
  return WriteAll(f, buf.data(), buf.size());
}

}  // namespace

extern "C" {

csg_status csg_camera_move(csg_camera* cam, float forward, float right, float up) {
  if (!cam) return Fail(CSG_E_ARG, "csg_camera_move: NULL camera");
  const csgray::CameraBasis b = csgray::MakeBasis(cam->pos, cam->pitch, cam->yaw, cam->vfov);
  cam->pos[0] += b.forward[0] * forward + b.right[0] * right;
  cam->pos[1] += b.forward[1] * forward + b.right[1] * right + up;
  cam->pos[2] += b.forward[2] * forward + b.right[2] * right;
  return CSG_OK;
}

csg_status csg_camera_rotate(csg_camera* cam, float dpitch, float dyaw) {
  if (!cam) return Fail(CSG_E_ARG, "csg_camera_rotate: NULL camera");
  const float lim = csgray::PitchLimit();
  cam->pitch = std::fmax(-lim, std::fmin(lim, cam->pitch + dpitch));
  cam->yaw = cam->yaw + dyaw;
  return CSG_OK;
}

csg_status csg_camera_input(csg_camera* cam, uint32_t keys, int32_t mouse_pressed, float mouse_dx, float mouse_dy) {
  if (!cam) return Fail(CSG_E_ARG, "csg_camera_input: NULL camera");
  float f = 0, r = 0, u = 0;
  if (keys & CSG_KEY_FORWARD) f += 0.1f;
  if (keys & CSG_KEY_BACKWARD) f -= 0.1f;
  if (keys & CSG_KEY_LEFT) r -= 0.1f;
  if (keys & CSG_KEY_RIGHT) r += 0.1f;
  if (keys & CSG_KEY_UP) u += 0.1f;
  if (keys & CSG_KEY_DOWN) u -= 0.1f;
  if (mouse_pressed) {
    const float yaw = -0.005f * mouse_dx, pitch = -0.005f * mouse_dy;
    csg_camera_rotate(cam, pitch, yaw);
  }
  return csg_camera_move(cam, f, r, u);
}

csg_status csg_write_ppm(const char* path, const float* rgba, int32_t w, int32_t h) {
  csg_status st = CheckImage(rgba, w, h);
  if (st != CSG_OK) return st;
  File f;
  if ((st = OpenForWrite(path, &f)) != CSG_OK) return st;
  char head[64];
  const int n = std::snprintf(head, sizeof head, "P6\n%d %d\n255\n", w, h);
  if ((st = WriteAll(f.get(), head, static_cast<size_t>(n))) != CSG_OK) return st;
  std::vector<uint8_t> row(static_cast<size_t>(w) * 3);
  for (int32_t y = h - 1; y >= 0; y--) {
    const float* src = rgba + static_cast<size_t>(y) * w * 4;
    for (int32_t x = 0; x < w; x++)
      for (int c = 0; c < 3; c++) row[3 * x + c] = ToByte(src[4 * x + c]);
    if ((st = WriteAll(f.get(), row.data(), row.size())) != CSG_OK) return st;
  }
  return CSG_OK;
}

csg_status csg_write_png(const char* path, const float* rgba, int32_t w, int32_t h) {
  csg_status st = CheckImage(rgba, w, h);
  if (st != CSG_OK) return st;
  // filtered scanlines: filter byte 0 then RGBA8, top row first
  const size_t stride = static_cast<size_t>(w) * 4 + 1;
  std::vector<uint8_t> raw(stride * static_cast<size_t>(h));
  for (int32_t y = 0; y < h; y++) {
    const float* src = rgba + static_cast<size_t>(h - 1 - y) * w * 4;
    uint8_t* dst = raw.data() + static_cast<size_t>(y) * stride;
    dst[0] = 0;
    for (size_t i = 0; i < static_cast<size_t>(w) * 4; i++) dst[1 + i] = ToByte(src[i]);
  }
  // zlib stream (RFC 1950) of stored deflate blocks (RFC 1951 section 3.2.4)
  std::vector<uint8_t> z;
  z.reserve(raw.size() + raw.size() / 65535 * 5 + 16);
  z.push_back(0x78);
  z.push_back(0x01);
  size_t pos = 0;
  do {
    const size_t n = std::min<size_t>(65535, raw.size() - pos);
    const bool last = pos + n == raw.size();
    z.push_back(last ? 1 : 0);
    z.push_back(static_cast<uint8_t>(n));
    z.push_back(static_cast<uint8_t>(n >> 8));
    z.push_back(static_cast<uint8_t>(~n));
    z.push_back(static_cast<uint8_t>(~n >> 8));
    z.insert(z.end(), raw.begin() + static_cast<std::ptrdiff_t>(pos),
             raw.begin() + static_cast<std::ptrdiff_t>(pos + n));
    pos += n;
  } while (pos < raw.size());
  uint32_t a = 1, b = 0;  // Adler-32
  for (uint8_t c : raw) {
    a = (a + c) % 65521u;
    b = (b + a) % 65521u;
  }
  PutBe32(z, (b << 16) | a);

  File f;
  if ((st = OpenForWrite(path, &f)) != CSG_OK) return st;
  static const uint8_t sig[8] = {0x89, 'P', 'N', 'G', 0x0D, 0x0A, 0x1A, 0x0A};
  if ((st = WriteAll(f.get(), sig, 8)) != CSG_OK) return st;
  std::vector<uint8_t> ihdr;
  PutBe32(ihdr, static_cast<uint32_t>(w));
  PutBe32(ihdr, static_cast<uint32_t>(h));
  ihdr.push_back(8);  // bit depth
  ihdr.push_back(6);  // colour type RGBA
  ihdr.push_back(0);  // deflate
  ihdr.push_back(0);  // adaptive filtering
  ihdr.push_back(0);  // no interlace
  if ((st = WriteChunk(f.get(), "IHDR", ihdr)) != CSG_OK) return st;
  if ((st = WriteChunk(f.get(), "IDAT", z)) != CSG_OK) return st;
  return WriteChunk(f.get(), "IEND", {});
}

csg_status csg_write_raw(const char* path, const void* data, int32_t elem_bytes, int32_t w, int32_t h,
                         int32_t flip_rows) {
  csg_status st = CheckImage(data, w, h);
  if (st != CSG_OK) return st;
  if (elem_bytes <= 0 || elem_bytes > 64) return Fail(CSG_E_ARG, "csg_write_raw: elem_bytes out of range");
  File f;
  if ((st = OpenForWrite(path, &f)) != CSG_OK) return st;
  const size_t row = static_cast<size_t>(w) * static_cast<size_t>(elem_bytes);
  const uint8_t* p = static_cast<const uint8_t*>(data);
  for (int32_t y = 0; y < h; y++) {
    const int32_t src = flip_rows ? h - 1 - y : y;
    if ((st = WriteAll(f.get(), p + static_cast<size_t>(src) * row, row)) != CSG_OK) return st;
  }
  return CSG_OK;
}

}  // extern "C"
