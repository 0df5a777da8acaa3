// csgrender: headless front end of libcsgray.
//
// Renders a scene file (the reference's preorder text format) with the
// reference application's camera, light and shading model and writes the
// image (PNG or PPM, top row first), and optionally the raw depth (f32),
// primitive id (i32) and hit flags (u8) of every sample.  A script replays
// the interactive controls frame by frame (Application.cpp:85-119):
//
//   forward | backward | left | right | up | down   keys held this frame
//   mouse DX DY                                      mouse dragged by (DX, DY)
//
// one line per frame; each frame is written to --frames PATTERN (printf
// pattern with the frame number, e.g. fly_%04d.png).
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "csg.h"

namespace {

void Usage() {
  std::fprintf(stderr,
               "usage: csgrender SCENE.txt [options]\n"
               "  --size W H                 image size (default 640 360)\n"
               "  --eye X Y Z --target X Y Z camera looking at a point (with --fov)\n"
               "  --fov DEG                  vertical field of view (default 90)\n"
               "  --camera X Y Z PITCH YAW   camera position and rotation in radians\n"
               "  --light X Y Z              direction toward the light\n"
               "  --light-angles P A         light from polar/azimuth angles in degrees\n"
               "  --phong KA KD KS N         shading (default 0.2 0.8 0.7 30)\n"
               "  --samples N --seed S       jittered samples per pixel (image: their mean)\n"
               "  --mode fast|exact          arithmetic (default fast)\n"
               "  --out FILE.png|FILE.ppm    image (default out.png)\n"
               "  --depth F --id F --flags F raw per-sample arrays, render row order\n"
               "  --script FILE --frames PAT replay interactive input, one image per frame\n"
               "  --time N                   render N more times and report the rate\n");
}

bool Fail(const char* what) {
  std::fprintf(stderr, "csgrender: %s: %s\n", what, csg_last_error());
  return false;
}

#define HIP_OK(x)                                                              \
  do {                                                                         \
    const hipError_t e_ = (x);                                                 \
    if (e_ != hipSuccess) {                                                    \
      std::fprintf(stderr, "csgrender: %s: %s\n", #x, hipGetErrorString(e_)); \
      return false;                                                            \
    }                                                                          \
  } while (0)

struct Options {
  std::string scene, out = "out.png", depth, id, flags, script, frames;
  int w = 640, h = 360, samples = 1, timeReps = 0;
  uint32_t seed = 0;
  bool lookAt = false, haveCamera = false, haveLight = false, lightAngles = false;
  float eye[3] = {0, 0, 5}, target[3] = {0, 0, 0}, fovDeg = 90.0f;
  float camPos[3] = {0, 0, 5}, pitch = 0, yaw = 0;
  float light[3] = {0, 0, 0}, polar = 0, azimuth = 0;
  bool havePhong = false;
  float phong[4] = {0.2f, 0.8f, 0.7f, 30.0f};
  int mode = CSG_MODE_FAST;
};

bool Parse(int argc, char** argv, Options* o) {
  if (argc < 2) return false;
  o->scene = argv[1];
  auto need = [&](int i, int n) { return i + n < argc; };
  for (int i = 2; i < argc; i++) {
    const std::string a = argv[i];
    auto f = [&](int k) { return static_cast<float>(std::atof(argv[i + k])); };
    if (a == "--size" && need(i, 2)) {
      o->w = std::atoi(argv[i + 1]);
      o->h = std::atoi(argv[i + 2]);
      i += 2;
    } else if (a == "--eye" && need(i, 3)) {
      for (int k = 0; k < 3; k++) o->eye[k] = f(k + 1);
      o->lookAt = true;
      i += 3;
    } else if (a == "--target" && need(i, 3)) {
      for (int k = 0; k < 3; k++) o->target[k] = f(k + 1);
      o->lookAt = true;
      i += 3;
    } else if (a == "--fov" && need(i, 1)) {
      o->fovDeg = f(1);
      i += 1;
    } else if (a == "--camera" && need(i, 5)) {
      for (int k = 0; k < 3; k++) o->camPos[k] = f(k + 1);
      o->pitch = f(4);
      o->yaw = f(5);
      o->haveCamera = true;
      i += 5;
    } else if (a == "--light" && need(i, 3)) {
      for (int k = 0; k < 3; k++) o->light[k] = f(k + 1);
      o->haveLight = true;
      i += 3;
    } else if (a == "--light-angles" && need(i, 2)) {
      o->polar = f(1);
      o->azimuth = f(2);
      o->lightAngles = true;
      i += 2;
    } else if (a == "--phong" && need(i, 4)) {
      for (int k = 0; k < 4; k++) o->phong[k] = f(k + 1);
      o->havePhong = true;
      i += 4;
    } else if (a == "--samples" && need(i, 1)) {
      o->samples = std::atoi(argv[++i]);
    } else if (a == "--seed" && need(i, 1)) {
      o->seed = static_cast<uint32_t>(std::strtoul(argv[++i], nullptr, 10));
    } else if (a == "--mode" && need(i, 1)) {
      const std::string m = argv[++i];
      if (m == "fast") o->mode = CSG_MODE_FAST;
      else if (m == "exact") o->mode = CSG_MODE_EXACT;
      else return false;
    } else if (a == "--out" && need(i, 1)) {
      o->out = argv[++i];
    } else if (a == "--depth" && need(i, 1)) {
      o->depth = argv[++i];
    } else if (a == "--id" && need(i, 1)) {
      o->id = argv[++i];
    } else if (a == "--flags" && need(i, 1)) {
      o->flags = argv[++i];
    } else if (a == "--script" && need(i, 1)) {
      o->script = argv[++i];
    } else if (a == "--frames" && need(i, 1)) {
      o->frames = argv[++i];
    } else if (a == "--time" && need(i, 1)) {
      o->timeReps = std::atoi(argv[++i]);
    } else {
      std::fprintf(stderr, "csgrender: unknown or incomplete option %s\n", a.c_str());
      return false;
    }
  }
  return o->w > 1 && o->h > 1 && o->samples >= 1;
}

struct Frame {
  int w = 0, h = 0, spp = 1;
  size_t n = 0;  // samples
  float* rgba = nullptr;
  float* depth = nullptr;
  int32_t* id = nullptr;
  uint8_t* flags = nullptr;
  ~Frame() {
    (void)hipFree(rgba);
    (void)hipFree(depth);
    (void)hipFree(id);
    (void)hipFree(flags);
  }
};

bool Alloc(Frame* f, int w, int h, int spp) {
  f->w = w;
  f->h = h;
  f->spp = spp;
  f->n = static_cast<size_t>(w) * h * spp;
  HIP_OK(hipMalloc(&f->rgba, f->n * 16));
  HIP_OK(hipMalloc(&f->depth, f->n * 4));
  HIP_OK(hipMalloc(&f->id, f->n * 4));
  HIP_OK(hipMalloc(&f->flags, f->n));
  return true;
}

bool Render(const csg_scene* s, const csg_camera& cam, const csg_light& light, const csg_phong& ph,
            const Options& o, Frame* f) {
  csg_outputs out{f->rgba, f->depth, f->id, f->flags};
  csg_render_opts opts{};
  opts.mode = o.mode;
  opts.samples = o.samples;
  opts.seed = o.seed;
  if (csg_render_ex(s, &cam, &light, &ph, f->w, f->h, &out, &opts, nullptr) != CSG_OK) return Fail("render");
  if (csg_sync(nullptr) != CSG_OK) return Fail("sync");
  return true;
}

// Writes the image (mean over samples) and the requested raw arrays.
bool Write(const Frame& f, const std::string& image, const Options& o) {
  std::vector<float> rgba(f.n * 4);
  HIP_OK(hipMemcpy(rgba.data(), f.rgba, f.n * 16, hipMemcpyDeviceToHost));
  std::vector<float> img(static_cast<size_t>(f.w) * f.h * 4, 0.0f);
  for (size_t p = 0; p < static_cast<size_t>(f.w) * f.h; p++)
    for (int s = 0; s < f.spp; s++)
      for (int c = 0; c < 4; c++) img[4 * p + c] += rgba[4 * (p * f.spp + s) + c] / static_cast<float>(f.spp);
  const bool png = image.size() >= 4 && image.compare(image.size() - 4, 4, ".ppm") != 0;
  const csg_status st = png ? csg_write_png(image.c_str(), img.data(), f.w, f.h)
                            : csg_write_ppm(image.c_str(), img.data(), f.w, f.h);
  if (st != CSG_OK) return Fail("writing the image");
  auto dump = [&](const std::string& path, const void* dev, int bytes) {
    if (path.empty()) return true;
    std::vector<uint8_t> host(f.n * static_cast<size_t>(bytes));
    HIP_OK(hipMemcpy(host.data(), dev, host.size(), hipMemcpyDeviceToHost));
    if (csg_write_raw(path.c_str(), host.data(), bytes * f.spp, f.w, f.h, 0) != CSG_OK) return Fail("raw dump");
    return true;
  };
  return dump(o.depth, f.depth, 4) && dump(o.id, f.id, 4) && dump(o.flags, f.flags, 1);
}

bool Run(const Options& o) {
  std::ifstream in(o.scene);
  if (!in) {
    std::fprintf(stderr, "csgrender: cannot read %s\n", o.scene.c_str());
    return false;
  }
  std::stringstream text;
  text << in.rdbuf();
  const std::string t = text.str();
  csg_scene* scene = nullptr;
  char err[512] = {0};
  if (csg_scene_parse(t.data(), t.size(), &scene, err, sizeof err) != CSG_OK) {
    std::fprintf(stderr, "csgrender: cannot load tree: %s\n", err);
    return false;
  }
  csg_camera cam;
  csg_camera_default(&cam);
  if (o.lookAt) {
    if (csg_camera_look_at(o.eye, o.target, o.fovDeg, &cam) != CSG_OK) return Fail("camera");
  } else {
    if (o.haveCamera) {
      for (int k = 0; k < 3; k++) cam.pos[k] = o.camPos[k];
      cam.pitch = o.pitch;
      cam.yaw = o.yaw;
    }
    cam.vfov = o.fovDeg * 3.14159f / 180.0f;
  }
  csg_light light;
  csg_light_default(&light);
  if (o.lightAngles) csg_light_from_angles(o.polar * 3.14159f / 180.0f, o.azimuth * 3.14159f / 180.0f, &light);
  if (o.haveLight)
    for (int k = 0; k < 3; k++) light.dir[k] = o.light[k];
  csg_phong ph;
  csg_phong_default(&ph);
  if (o.havePhong) {
    ph.ka = o.phong[0];
    ph.kd = o.phong[1];
    ph.ks = o.phong[2];
    ph.shininess = o.phong[3];
  }
  Frame f;
  bool ok = Alloc(&f, o.w, o.h, o.samples);
  if (ok && o.script.empty()) {
    ok = Render(scene, cam, light, ph, o, &f) && Write(f, o.out, o);
    if (ok && o.timeReps > 0) {
      const auto t0 = std::chrono::steady_clock::now();
      for (int r = 0; ok && r < o.timeReps; r++) ok = Render(scene, cam, light, ph, o, &f);
      const double ms =
          std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count() / o.timeReps;
      if (ok)
        std::printf("%.3f ms/frame, %.1f Mrays/s (%d x %d x %d)\n", ms, static_cast<double>(f.n) / ms / 1e3, o.w, o.h,
                    o.samples);
    }
  } else if (ok) {
    std::ifstream sc(o.script);
    if (!sc) {
      std::fprintf(stderr, "csgrender: cannot read %s\n", o.script.c_str());
      ok = false;
    }
    const std::string pattern = o.frames.empty() ? "frame_%04d.png" : o.frames;
    std::string line;
    for (int frame = 0; ok && std::getline(sc, line); frame++) {
      std::istringstream ls(line);
      std::string tok;
      uint32_t keys = 0;
      int pressed = 0;
      float dx = 0, dy = 0;
      while (ok && ls >> tok) {
        if (tok == "forward") keys |= CSG_KEY_FORWARD;
        else if (tok == "backward") keys |= CSG_KEY_BACKWARD;
        else if (tok == "left") keys |= CSG_KEY_LEFT;
        else if (tok == "right") keys |= CSG_KEY_RIGHT;
        else if (tok == "up") keys |= CSG_KEY_UP;
        else if (tok == "down") keys |= CSG_KEY_DOWN;
        else if (tok == "mouse" && (ls >> dx >> dy)) pressed = 1;
        else {
          std::fprintf(stderr, "csgrender: script line %d: bad token %s\n", frame + 1, tok.c_str());
          ok = false;
        }
      }
      if (!ok) break;
      csg_camera_input(&cam, keys, pressed, dx, dy);
      char name[1024];
      std::snprintf(name, sizeof name, pattern.c_str(), frame);
      ok = Render(scene, cam, light, ph, o, &f) && Write(f, name, o);
    }
  }
  csg_scene_free(scene);
  return ok;
}

}  // namespace

int main(int argc, char** argv) {
  Options o;
  if (!Parse(argc, argv, &o)) {
    Usage();
    return 2;
  }
  return Run(o) ? 0 : 1;
}
