/* csg.h — C ABI of the MI355X CSG ray caster (libcsgray.so).
 *
 * The reference exposes its render path as two C++ entry points, and this
 * header is the drop-in for them:
 *   CSGTree::Parse(const std::string&)          Graphics/RayCasting/CSGTree/CSGTree.cuh:53,
 *                                               CSGTree.cu:5-152 (parse + ConstructBVH)
 *   Raycaster::ChangeSize(int, int, CSGTree)    Graphics/RayCasting/Raycaster.cuh:21, Raycaster.cu:3-21
 *   Raycaster::Raycast(float4*, Camera, DirectionalLight)
 *                                               Raycaster.cuh:22, Raycaster.cu:23-34
 *   Raycaster::CleanUp()                        Raycaster.cuh:23, Raycaster.cu:36-45
 * plus the value types it passes by value: Camera (RenderManager/Camera/Camera.h:3-54)
 * and DirectionalLight (RenderManager/DirectionalLight.h:8-18).
 *
 * Conventions
 *   - Every function returns a csg_status; none throws or exits.  The text of
 *     the last error on the calling thread is available from csg_last_error().
 *   - Scenes are immutable after csg_scene_parse and may be rendered
 *     concurrently from several host threads and streams.
 *   - Output pointers are caller-owned device pointers (any may be NULL).  The
 *     image is row-major, pixel (x, y) at index y*width + x; row 0 is the
 *     bottom scanline when displayed (its rays point down), as in the reference.
 *   - `stream` is a hipStream_t passed as void* (NULL = the null stream).  All
 *     renders are stream-ordered and asynchronous; csg_sync waits.
 */
#ifndef CSG_H_
#define CSG_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum csg_status {
  CSG_OK = 0,
  CSG_E_PARSE = 1, /* scene text rejected (the reference throws std::invalid_argument) */
  CSG_E_DEPTH = 2, /* operator depth above what the traversal stack holds */
  CSG_E_OOM = 3,   /* device or host allocation failed */
  CSG_E_HIP = 4,   /* a HIP runtime call failed */
  CSG_E_ARG = 5,   /* invalid argument (null pointer, size, range) */
  CSG_E_BUDGET = 6 /* serving: request larger than the admission budget */
} csg_status;

typedef struct csg_scene csg_scene;

/* Camera: the reference's pitch/yaw model (Camera.h:9-16).  pitch = rotX,
 * yaw = rotY, both in radians; vfov = vertical field of view in radians.
 * The basis is forward = (-sin yaw cos pitch, sin pitch, -cos yaw cos pitch),
 * right = (cos yaw, 0, -sin yaw), up = forward x right (Camera.cpp:4-27);
 * pitch is clamped to +-89 degrees as Camera::setRotation does. */
typedef struct csg_camera {
  float pos[3];
  float pitch;
  float yaw;
  float vfov;
} csg_camera;

/* Directional light: `dir` points TOWARD the light, as the reference's
 * getLightDir() result is used (RaycastingKernels.cu:78-86).  It need not be
 * normalised. */
typedef struct csg_light {
  float dir[3];
} csg_light;

/* Phong constants.  The reference hard-codes ka 0.2, kd 0.8, ks 0.7,
 * shininess 30 and background (0.08, 0.08, 0.11, 1)
 * (RaycastingKernels.cu:71-74,109); csg_phong_default returns those. */
typedef struct csg_phong {
  float ka, kd, ks, shininess;
  float bg[4];
} csg_phong;

/* Per-pixel outputs (device pointers, each may be NULL):
 *   rgba    float4 per pixel, the reference's RGBA32F framebuffer
 *   depth   hit distance t along the unit ray (+inf on a miss)
 *   prim_id index of the visible primitive in file order (-1 on a miss)
 *   flags   the final hit bits: 2 enter, 4 exit, 8 miss, 16 flipped,
 *           32 bottom cap, 64 top cap (Utils/CSGUtils.cuh:27-35) */
typedef struct csg_outputs {
  float* rgba;
  float* depth;
  int32_t* prim_id;
  uint8_t* flags;
} csg_outputs;

typedef struct csg_scene_info {
  int32_t nodes;
  int32_t primitives;
  int32_t op_depth;  /* operator levels above the deepest leaf */
  int32_t spheres, cylinders, cubes;
  int64_t device_bytes;
  int32_t prims_in_boxes;  /* 1: every primitive lies inside its box (a cylinder tilted off the
                              coordinate axes may not: the reference's box is centre +- max(h/2, r)),
                              so handed-over rays may be classified by intervals */
  int32_t union_groups;    /* Union subtrees (3 to 1024 members below their Unions) that interval
                              classification walks as one operator, testing all members' boxes at once */
} csg_scene_info;

/* Arithmetic of the traversal kernel.
 *   CSG_MODE_EXACT: the reference's FP32 expressions as written — IEEE
 *     division and square root, no fused multiply-add — so hit id, depth and
 *     flags are bit-identical to a host FP32 build of the reference logic.
 *   CSG_MODE_FAST: box slabs as one fused multiply-add each from a per-ray
 *     reciprocal direction, sphere and cylinder quadratics from the ray's
 *     closest approach to the centre or axis (better conditioned than the
 *     reference's expressions), hardware square roots and reciprocals where
 *     they do not move a result; decisions stay the reference's and results
 *     are checked against the FP64 host oracle (depth within 1e-4). */
typedef enum csg_mode { CSG_MODE_EXACT = 0, CSG_MODE_FAST = 1 } csg_mode;

typedef struct csg_render_opts {
  int32_t mode;         /* csg_mode */
  int32_t block;        /* threads per workgroup: 0 = auto, else 256, 512 or 1024 */
  int32_t scene_placement; /* 0 = auto, 1 = read the scene from global memory, 2 = stage it in LDS,
                              3 = global memory with the top of the tree (as many operator records,
                              in level order, as fit beside the stacks) staged in LDS: for scenes
                              too large for LDS; single-scene renders only */
  int32_t counters;     /* 1: also write per-ray event counters (testing) */
  int32_t* counter_out; /* device, 6 int32 per pixel: GoTo, Compute, box, sphere, cylinder, cube */
  const uint32_t* tile_order; /* device, optional: order in which the ceil(w/8) x ceil(h/8) tiles of
                                 8x8 pixels (row-major ids) are taken; a permutation, or NULL */
  int32_t samples;      /* rays per pixel (0 or 1: the reference's single ray through the pixel centre);
                           with n > 1 every output buffer holds w*h*n entries, sample s of pixel p at
                           p*n + s, each ray offset within its pixel by csg_jitter() */
  uint32_t seed;        /* jitter seed */
  int32_t tile_walk;    /* 0: row-major tiles; 1: Morton order within a job (square power-of-two grids) */
  int32_t coop_budget;  /* state-machine steps a lane spends on one ray before the ray may be handed to a
                           whole wave (cooperative evaluation, same results): 0 = default (384; 64 where
                           the wave classifies by intervals, see coop_intervals), < 0 = never */
  int32_t coop_when;    /* 0: hand over only in the frame's tail, once every tile has been taken and a
                           further coop_grace percent of the time that took has passed;
                           1: as soon as the budget is spent */
  int32_t coop_grace;   /* 0 = default: 950 x waves / tiles percent (at most 300) for one scene, 550 x for
                           a batch (the tiles in flight when the queue runs dry finish in about a tile's
                           time; handing out took tiles / waves tile times), and 95000 x (waves /
                           tiles)^5 (0.5 to 100) for one scene whose rays the wave classifies by
                           intervals (such rays cost the wave little: the tail goes over at once
                           unless the frame has few tiles a wave): 120 (3 with intervals) for a
                           1024-lane launch of 1920x1080, about 3 for the 256-tree batch; < 0 = none;
                           else percent */
  int32_t coop_intervals; /* 0 = default: in fast mode, a handed-over ray of a scene whose primitives all
                             lie inside their boxes is classified by intervals (one pass over the hit part
                             of the tree, the crossings sorted, membership masks combined bottom-up; the
                             state machine's result, which a ray gets whenever the two could order events
                             differently); < 0 = state machine only */
  int32_t tight_boxes;  /* 0 = default: fast mode culls with boxes tight around each subtree's solid
                           where the reference's culling allows (a sphere's c +- r where the reference
                           builds c +- 2r; a cylinder's own extent unless it pokes out of the
                           reference's box, which it then keeps; a Difference's left operand's box; an
                           Intersection's overlap): the same results on every ray whose culling the
                           reference does not decide within a rounding margin, in fewer steps;
                           < 0 = the reference's boxes */
} csg_render_opts;

/* --- scene (replaces CSGTree::Parse + the upload in Raycaster::ChangeSize) --- */
/* Parse, build boxes, flatten and upload to the current device. */
csg_status csg_scene_parse(const char* text, size_t len, csg_scene** out, char* err, size_t errlen);
/* The host half alone (CSGTree::Parse + ConstructBVH): no device memory is
 * touched; csg_scene_upload later copies the scene to the current device. */
csg_status csg_scene_parse_host(const char* text, size_t len, csg_scene** out, char* err, size_t errlen);
csg_status csg_scene_upload(csg_scene* scene);
csg_status csg_scene_free(csg_scene* scene);
csg_status csg_scene_info_get(const csg_scene* scene, csg_scene_info* info);
/* Host copy of the flattened scene, for tests: links = 5 int32 per node
 * (type, primitive, left, right, parent), boxes = 6 float per node
 * (min xyz, max xyz), prims = 14 float per primitive (id, type, x, y, z,
 * r, g, b, radius, height, size, axis xyz).  Any pointer may be NULL. */
csg_status csg_scene_export(const csg_scene* scene, int32_t* links, float* boxes, float* prims);

/* --- camera / light / shading helpers --- */
void csg_camera_default(csg_camera* cam);           /* Camera() : (0,0,5), pitch=yaw=0, fov 90 deg */
/* eye -> target, vfov in degrees converted with 3.14159 as Camera::setFOV does */
csg_status csg_camera_look_at(const float eye[3], const float target[3], float vfov_deg, csg_camera* cam);
/* forward[3], right[3], up[3], tan(vfov/2) as the render computes them */
csg_status csg_camera_basis(const csg_camera* cam, float out10[10]);
void csg_light_default(csg_light* light);           /* polar -60 deg, azimuth -45 deg */
void csg_light_from_angles(float polar, float azimuth, csg_light* light);
void csg_phong_default(csg_phong* phong);

/* Camera control of the interactive application, for scripted fly-throughs.
 * csg_camera_move: Camera::move (Camera.h:37-42), along the current forward
 * and right vectors, `up` along world Y.  csg_camera_rotate: Camera::rotate
 * (Camera.h:44-47), pitch clamped to +-89 degrees as in setRotation.
 * csg_camera_input: one frame of Application::Input (Application.cpp:85-119):
 * keys step 0.1 (forward/backward, right/left, up/down); with the mouse
 * pressed, yaw -= 0.005 * dx and pitch -= 0.005 * dy, applied before the move. */
enum {
  CSG_KEY_FORWARD = 1, CSG_KEY_BACKWARD = 2, CSG_KEY_LEFT = 4, CSG_KEY_RIGHT = 8,
  CSG_KEY_UP = 16, CSG_KEY_DOWN = 32
};
csg_status csg_camera_move(csg_camera* cam, float forward, float right, float up);
csg_status csg_camera_rotate(csg_camera* cam, float dpitch, float dyaw);
csg_status csg_camera_input(csg_camera* cam, uint32_t keys, int32_t mouse_pressed, float mouse_dx, float mouse_dy);

/* Headless output (no reference counterpart: the reference only displays
 * through OpenGL).  Inputs are HOST buffers in render order (row 0 = the
 * bottom scanline); image files are written top row first, i.e. rows are
 * flipped.  Channels map [0,1] to 0..255 (round to nearest, clamped, NaN 0).
 * csg_write_ppm: binary P6 (RGB).  csg_write_png: 8-bit RGBA, stored
 * (uncompressed) deflate blocks.  csg_write_raw: the array as is (little
 * endian, `elem_bytes` per pixel), rows flipped when `flip_rows`. */
csg_status csg_write_ppm(const char* path, const float* rgba, int32_t width, int32_t height);
csg_status csg_write_png(const char* path, const float* rgba, int32_t width, int32_t height);
csg_status csg_write_raw(const char* path, const void* data, int32_t elem_bytes, int32_t width, int32_t height,
                         int32_t flip_rows);

/* --- render (replaces Raycaster::Raycast: RaycastKernel + LightningKernel) --- */
csg_status csg_render(const csg_scene* scene, const csg_camera* cam, const csg_light* light,
                      const csg_phong* phong, int32_t width, int32_t height, const csg_outputs* out,
                      void* stream);
csg_status csg_render_ex(const csg_scene* scene, const csg_camera* cam, const csg_light* light,
                         const csg_phong* phong, int32_t width, int32_t height, const csg_outputs* out,
                         const csg_render_opts* opts, void* stream);
/* csg_render_ex that allocates nothing: no counterpart in the reference;
 * for callers that cannot allocate device memory (code under stream capture,
 * tight serving loops).  Same arguments and results as csg_render_ex.  Its
 * scratch (work-queue words, a heavy-ray list of 4,194,304 entries and a
 * 16 MiB resume area, about 48 MiB) is one static device array per GPU,
 * held until the process ends; calls of more than 4,194,304 rays
 * (width x height x samples) fail with CSG_E_ARG.  Every call on one GPU
 * shares that array, so calls that overlap on one GPU (several streams or
 * host threads) race: serialise them. */
csg_status csg_render_noalloc(const csg_scene* scene, const csg_camera* cam, const csg_light* light,
                              const csg_phong* phong, int32_t width, int32_t height, const csg_outputs* out,
                              const csg_render_opts* opts, void* stream);
csg_status csg_sync(void* stream);
/* Each (device, stream) pair gets a small workspace (tile queue, heavy-ray
 * list) on its first render, cleared with hipMemsetAsync on that stream; the
 * kernels leave it cleared after every launch.  To capture csg_render /
 * csg_render_ex / csg_render_batch into a graph, render once on the stream
 * first, so that no memset is recorded. */

/* Rays handed to the cooperative kernel so far by csg_render, csg_render_ex
 * and csg_render_batch launches on `stream` (current device); waits for the
 * stream.  No reference counterpart: for tuning the hand-over and for tests. */
csg_status csg_coop_rays(void* stream, int64_t* rays);
/* Of those, the rays the cooperative kernel finished by interval
 * classification (csg_render_opts.coop_intervals) rather than with the state
 * machine. */
csg_status csg_coop_interval_rays(void* stream, int64_t* rays);

/* Sub-pixel offset in [0,1) of sample `sample` of pixel `pixel` along axis 0
 * (x) or 1 (y): a 32-bit integer hash of (seed, pixel, sample, axis) with 24
 * bits kept; the kernels use the same function.  (No reference counterpart:
 * the reference casts one ray per pixel through its centre.) */
float csg_jitter(uint32_t seed, uint32_t pixel, uint32_t sample, uint32_t axis);

/* --- batched render (no reference counterpart: the reference renders one
 * tree per launch).  One persistent launch renders every job; waves take
 * 8x8 tiles from one queue across all jobs, in the order the jobs are given
 * (csg_batch_order sorts them heaviest first). --- */
typedef struct csg_job {
  const csg_scene* scene;
  csg_camera camera;
  csg_light light;
  csg_phong phong;
  int32_t width, height;
  csg_outputs out;
} csg_job;
csg_status csg_render_batch(const csg_job* jobs, int32_t njobs, const csg_render_opts* opts, void* stream);
/* Fills order[njobs] with job indices by descending predicted cost
 * (pixels x (1 + operator depth)). */
csg_status csg_batch_order(const csg_job* jobs, int32_t njobs, int32_t* order);

/* --- captured batches (no reference counterpart): the batch's job tables and
 * tile queues live in its own device buffers and its launches are captured
 * once in a hipGraph; the steady state is csg_batch_launch. --- */
typedef struct csg_batch csg_batch;
csg_status csg_batch_create(const csg_job* jobs, int32_t njobs, const csg_render_opts* opts, csg_batch** out);
/* New jobs (cameras, lights, outputs, sizes; the count may change only
 * within the captured number of launches): the device job tables are
 * rewritten in order on `stream`, and when tile counts or stack depth change
 * the graph's kernel nodes get new parameters. */
csg_status csg_batch_update(csg_batch* batch, const csg_job* jobs, int32_t njobs, void* stream);
csg_status csg_batch_launch(csg_batch* batch, void* stream);
/* Times candidate launch shapes on the batch (workgroup 256/512/1024 lanes x
 * row-major/Morton tile walk x given/heaviest-first job order x LDS stack
 * depth: one launch whose stacks fit the deepest tree, or two depth classes
 * split at the depth that halves the batch's work (pixels x depth; failing
 * that, below the deepest trees), each launch's stacks sized to its own
 * deepest tree; `reps` launches each), keeps the fastest for later launches
 * and reports it. */
typedef struct csg_batch_tuning {
  int32_t block;
  int32_t tile_walk;
  int32_t heaviest_first;
  float ms;  /* mean time of one launch of the batch with this choice */
  int32_t depth_split; /* 0: one launch; d > 0: trees of <= d levels launch apart from deeper ones */
} csg_batch_tuning;
csg_status csg_batch_autotune(csg_batch* batch, int32_t reps, void* stream, csg_batch_tuning* chosen);
csg_status csg_batch_free(csg_batch* batch);

/* Launch geometry csg_render_ex would use for this scene and options on the
 * current device, for a large frame (no reference counterpart; for tests and
 * tuning).  A single-scene launch with the scene in LDS takes 1024-lane
 * workgroups when they fit, the scene has at least 256 operators and the frame
 * has an 8x8 tile for every wave of that grid; otherwise 512 lanes. */
typedef struct csg_launch_info {
  int32_t block;          /* threads per workgroup */
  int32_t grid;           /* workgroups (persistent: waves pull 8x8 tiles from a queue) */
  int64_t lds_bytes;      /* dynamic LDS per workgroup */
  int32_t scene_in_lds;   /* 1 if the scene is staged in LDS */
  int32_t compute_units;
  int64_t lds_limit;      /* LDS per workgroup the device allows */
  int32_t coop_budget;    /* steps before a ray is handed to the cooperative kernel (0: never) */
  int32_t coop_grid;      /* workgroups of the cooperative kernel (one ray in flight each) */
  int32_t coop_block;     /* its threads per workgroup (64) */
  int32_t coop_grace;    /* tail grace (percent) a 1920x1080 frame would get (the default of
                            csg_render_opts.coop_grace) */
  int64_t coop_lds_bytes; /* its dynamic LDS per workgroup */
  int32_t cache_ops;      /* operator records cached in LDS (scene_placement 3), else 0 */
} csg_launch_info;
csg_status csg_render_plan(const csg_scene* scene, const csg_render_opts* opts, csg_launch_info* info);

/* --- serving (no reference counterpart: the reference renders one frame at a
 * time for one window).  A server keeps scenes resident, queues render
 * requests, admits them in arrival order while they fit a device memory
 * budget, and keeps up to `streams` batches in flight: requests that arrive
 * while others run join the next batch on a free stream (continuous
 * batching).  Outputs are server-owned until released.
 *
 * The budget counts all device memory the server holds: each stream's
 * workspace (tile queue, heavy-ray list and resume area, job table; about
 * 96 MiB a stream, allocated by csg_server_create, which fails with
 * CSG_E_BUDGET when the budget does not cover them) and the output blocks
 * of admitted requests until they are released.  Released blocks are reused
 * by later requests; a block copied out with csg_server_copy is not written
 * again before that copy has run on the caller's stream, so a request may be
 * released right after its copy is enqueued. --- */
typedef struct csg_server csg_server;
typedef struct csg_server_config {
  int64_t budget_bytes;   /* device memory for request outputs; 0 = 16 GiB */
  int32_t streams;        /* batches in flight; 0 = 3 */
  int32_t max_batch_jobs; /* requests per batch; 0 = 256 */
  int32_t mode;           /* csg_mode */
  int32_t coop_budget;    /* as csg_render_opts.coop_budget */
  int32_t coop_when;      /* as csg_render_opts.coop_when */
  int32_t coop_grace;     /* as csg_render_opts.coop_grace */
} csg_server_config;
typedef struct csg_request {
  int32_t scene_id;
  csg_camera camera;
  csg_light light;
  csg_phong phong;
  int32_t width, height;
} csg_request;
typedef struct csg_request_stats {
  double submit_s, start_s, done_s;  /* seconds since the server was created */
  int32_t batch_jobs;                /* requests in the batch it ran in */
  int32_t status;                    /* csg_status of its launch (outputs undefined unless CSG_OK) */
} csg_request_stats;
typedef struct csg_server_stats {
  int64_t submitted, completed, rejected, batches, rays_done;
  int64_t bytes_in_use, peak_bytes_in_use; /* workspaces + admitted outputs */
  int64_t workspace_bytes;                 /* the streams' workspaces */
} csg_server_stats;
csg_status csg_server_create(const csg_server_config* cfg, csg_server** out);
/* The scene must stay alive and on the server's device while the server runs. */
csg_status csg_server_add_scene(csg_server* server, const csg_scene* scene, int32_t* scene_id);
/* CSG_E_BUDGET when the request's outputs exceed the budget left beside the
 * streams' workspaces.  A request whose job fails validation when its batch
 * forms is done at once with that status (csg_request_stats.status); one
 * whose launch fails is done with CSG_E_HIP. */
csg_status csg_server_submit(csg_server* server, const csg_request* req, int64_t* ticket);
csg_status csg_server_poll(csg_server* server, int64_t ticket, int32_t* done);
/* Blocks until the request is done; fills the server-owned device output
 * pointers (valid until release) and the request's timing. */
csg_status csg_server_wait(csg_server* server, int64_t ticket, csg_outputs* out, csg_request_stats* stats);
/* Copies a done request's outputs into caller-owned device buffers (any may
 * be NULL), asynchronously on `stream`; the request may be released at once. */
csg_status csg_server_copy(csg_server* server, int64_t ticket, const csg_outputs* dst, void* stream);
csg_status csg_server_release(csg_server* server, int64_t ticket);
csg_status csg_server_stats_get(csg_server* server, csg_server_stats* stats);
/* Waits for every request, then frees the server's memory and streams. */
csg_status csg_server_destroy(csg_server* server);

const char* csg_last_error(void);
const char* csg_version(void);

#ifdef __cplusplus
}
#endif

#endif /* CSG_H_ */
