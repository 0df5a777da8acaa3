// Internals shared by the C ABI's translation units: the scene handle,
// error reporting, option parsing, job construction and per-stream device
// workspaces.
#pragma once

#include <hip/hip_runtime.h>

#include <string>

#include "csg.h"
#include "device.hpp"
#include "scene.hpp"

struct csg_scene {
  csgray::CSGTree tree;
  csgray::FlatScene flat;
  int device = 0;
  void* dmem = nullptr;
  size_t bytes = 0;
  csgray::DevScene dev{};
};

namespace csgray {
namespace rt {

// Records `msg` as the calling thread's last error and returns `s`.
csg_status Fail(csg_status s, const std::string& msg);
csg_status FailHip(hipError_t e, const char* what);

csg_status ParseOpts(const csg_render_opts* opts, LaunchConfig* cfg);
size_t SceneBytes(const csg_scene* s);

// Validates one job and fills its device descriptor; `tileBase` is its first
// queue index within the launch.
csg_status BuildJob(const csg_scene* s, const csg_camera* cam, const csg_light* light, const csg_phong* phong,
                    int32_t width, int32_t height, const csg_outputs* out, const csg_render_opts* opts, int dev,
                    int64_t tileBase, Job* jb);

// The tile queue (five words, zero between launches) of (device, stream).
csg_status QueueFor(int dev, hipStream_t stream, uint32_t** out);

// The heavy-ray list of (device, stream).
csg_status HeavyFor(int dev, hipStream_t stream, uint2** out);

// Fills the cooperative-evaluation fields of `p` for a launch whose scenes
// have at most `maxOps` operators (deferral off when the options say so or
// the cooperative kernel's state does not fit LDS).
void PrepareCoop(const LaunchConfig& cfg, int32_t maxOps, int32_t maxLeaves, bool anyInBoxes, uint2* heavy,
                 const LaunchShape& shape, RenderParams* p);

// Copies a job table to the (device, stream) table buffer in stream order.
csg_status UploadJobs(int dev, hipStream_t stream, const Job* jobs, size_t n, const Job** out);

// Device bytes of one stream's workspace (queue, heavy-ray list with its
// resume area, job table for up to `maxJobs` jobs).
size_t StreamWorkspaceBytes(int32_t maxJobs);

// Frees the queue, heavy-ray list and job table of (device, stream); the
// caller has synchronised the stream and renders no more on it.
csg_status ReleaseStreamWorkspace(int dev, hipStream_t stream);

}  // namespace rt
}  // namespace csgray
