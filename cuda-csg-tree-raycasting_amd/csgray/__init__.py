"""ctypes binding of libcsgray.so (the C ABI in include/csg.h).

This is plumbing for tests, the benchmark and scripted renders: every call
goes straight to the native library.  Device buffers come from torch (its
allocator and its current stream); nothing here computes pixels.  If the
library has not been built, `lib()` raises — there is no fallback.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass

_PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(_PKG_DIR, "lib", "libcsgray.so")

c_float3 = ctypes.c_float * 3

# csg_status
OK, E_PARSE, E_DEPTH, E_OOM, E_HIP, E_ARG, E_BUDGET = range(7)
STATUS_NAMES = {0: "CSG_OK", 1: "CSG_E_PARSE", 2: "CSG_E_DEPTH", 3: "CSG_E_OOM", 4: "CSG_E_HIP",
                5: "CSG_E_ARG", 6: "CSG_E_BUDGET"}
MODE_EXACT, MODE_FAST = 0, 1
PLACE_AUTO, PLACE_GLOBAL, PLACE_LDS, PLACE_CACHE = 0, 1, 2, 3

# hit flag bits (final record)
ENTER, EXIT, MISS, FLIP, CAP_BOTTOM, CAP_TOP = 2, 4, 8, 16, 32, 64


class Camera(ctypes.Structure):
    _fields_ = [("pos", ctypes.c_float * 3), ("pitch", ctypes.c_float), ("yaw", ctypes.c_float),
                ("vfov", ctypes.c_float)]


class Light(ctypes.Structure):
    _fields_ = [("dir", ctypes.c_float * 3)]


class Phong(ctypes.Structure):
    _fields_ = [("ka", ctypes.c_float), ("kd", ctypes.c_float), ("ks", ctypes.c_float),
                ("shininess", ctypes.c_float), ("bg", ctypes.c_float * 4)]


class Outputs(ctypes.Structure):
    _fields_ = [("rgba", ctypes.c_void_p), ("depth", ctypes.c_void_p), ("prim_id", ctypes.c_void_p),
                ("flags", ctypes.c_void_p)]


class SceneInfo(ctypes.Structure):
    _fields_ = [("nodes", ctypes.c_int32), ("primitives", ctypes.c_int32), ("op_depth", ctypes.c_int32),
                ("spheres", ctypes.c_int32), ("cylinders", ctypes.c_int32), ("cubes", ctypes.c_int32),
                ("device_bytes", ctypes.c_int64), ("prims_in_boxes", ctypes.c_int32),
                ("union_groups", ctypes.c_int32)]


class RenderOpts(ctypes.Structure):
    _fields_ = [("mode", ctypes.c_int32), ("block", ctypes.c_int32), ("scene_placement", ctypes.c_int32),
                ("counters", ctypes.c_int32), ("counter_out", ctypes.c_void_p), ("tile_order", ctypes.c_void_p),
                ("samples", ctypes.c_int32), ("seed", ctypes.c_uint32), ("tile_walk", ctypes.c_int32),
                ("coop_budget", ctypes.c_int32), ("coop_when", ctypes.c_int32), ("coop_grace", ctypes.c_int32),
                ("coop_intervals", ctypes.c_int32), ("tight_boxes", ctypes.c_int32)]


class Job(ctypes.Structure):
    _fields_ = [("scene", ctypes.c_void_p), ("camera", Camera), ("light", Light), ("phong", Phong),
                ("width", ctypes.c_int32), ("height", ctypes.c_int32), ("out", Outputs)]


class LaunchInfo(ctypes.Structure):
    _fields_ = [("block", ctypes.c_int32), ("grid", ctypes.c_int32), ("lds_bytes", ctypes.c_int64),
                ("scene_in_lds", ctypes.c_int32), ("compute_units", ctypes.c_int32), ("lds_limit", ctypes.c_int64),
                ("coop_budget", ctypes.c_int32), ("coop_grid", ctypes.c_int32), ("coop_block", ctypes.c_int32),
                ("coop_grace", ctypes.c_int32), ("coop_lds_bytes", ctypes.c_int64), ("cache_ops", ctypes.c_int32)]


class BatchTuning(ctypes.Structure):
    _fields_ = [("block", ctypes.c_int32), ("tile_walk", ctypes.c_int32), ("heaviest_first", ctypes.c_int32),
                ("ms", ctypes.c_float), ("depth_split", ctypes.c_int32)]


class ServerConfig(ctypes.Structure):
    _fields_ = [("budget_bytes", ctypes.c_int64), ("streams", ctypes.c_int32), ("max_batch_jobs", ctypes.c_int32),
                ("mode", ctypes.c_int32), ("coop_budget", ctypes.c_int32), ("coop_when", ctypes.c_int32),
                ("coop_grace", ctypes.c_int32)]


class Request(ctypes.Structure):
    _fields_ = [("scene_id", ctypes.c_int32), ("camera", Camera), ("light", Light), ("phong", Phong),
                ("width", ctypes.c_int32), ("height", ctypes.c_int32)]


class RequestStats(ctypes.Structure):
    _fields_ = [("submit_s", ctypes.c_double), ("start_s", ctypes.c_double), ("done_s", ctypes.c_double),
                ("batch_jobs", ctypes.c_int32), ("status", ctypes.c_int32)]


class ServerStats(ctypes.Structure):
    _fields_ = [("submitted", ctypes.c_int64), ("completed", ctypes.c_int64), ("rejected", ctypes.c_int64),
                ("batches", ctypes.c_int64), ("rays_done", ctypes.c_int64), ("bytes_in_use", ctypes.c_int64),
                ("peak_bytes_in_use", ctypes.c_int64), ("workspace_bytes", ctypes.c_int64)]


class CsgError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {msg}")
        self.status = status


_LIB = None

# (name, restype, argtypes) for every function declared in include/csg.h
_P = ctypes.c_void_p
_SIGNATURES = [
    ("csg_scene_parse", ctypes.c_int, [ctypes.c_char_p, ctypes.c_size_t, ctypes.POINTER(_P), ctypes.c_char_p,
                                       ctypes.c_size_t]),
    ("csg_scene_parse_host", ctypes.c_int, [ctypes.c_char_p, ctypes.c_size_t, ctypes.POINTER(_P),
                                            ctypes.c_char_p, ctypes.c_size_t]),
    ("csg_scene_upload", ctypes.c_int, [_P]),
    ("csg_scene_free", ctypes.c_int, [_P]),
    ("csg_scene_info_get", ctypes.c_int, [_P, ctypes.POINTER(SceneInfo)]),
    ("csg_scene_export", ctypes.c_int, [_P, _P, _P, _P]),
    ("csg_camera_default", None, [ctypes.POINTER(Camera)]),
    ("csg_camera_look_at", ctypes.c_int, [c_float3, c_float3, ctypes.c_float, ctypes.POINTER(Camera)]),
    ("csg_camera_basis", ctypes.c_int, [ctypes.POINTER(Camera), ctypes.c_float * 10]),
    ("csg_light_default", None, [ctypes.POINTER(Light)]),
    ("csg_light_from_angles", None, [ctypes.c_float, ctypes.c_float, ctypes.POINTER(Light)]),
    ("csg_phong_default", None, [ctypes.POINTER(Phong)]),
    ("csg_render", ctypes.c_int, [_P, ctypes.POINTER(Camera), ctypes.POINTER(Light), ctypes.POINTER(Phong),
                                  ctypes.c_int32, ctypes.c_int32, ctypes.POINTER(Outputs), _P]),
    ("csg_render_ex", ctypes.c_int, [_P, ctypes.POINTER(Camera), ctypes.POINTER(Light), ctypes.POINTER(Phong),
                                     ctypes.c_int32, ctypes.c_int32, ctypes.POINTER(Outputs),
                                     ctypes.POINTER(RenderOpts), _P]),
    ("csg_render_noalloc", ctypes.c_int, [_P, ctypes.POINTER(Camera), ctypes.POINTER(Light), ctypes.POINTER(Phong),
                                          ctypes.c_int32, ctypes.c_int32, ctypes.POINTER(Outputs),
                                          ctypes.POINTER(RenderOpts), _P]),
    ("csg_sync", ctypes.c_int, [_P]),
    ("csg_coop_rays", ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_int64)]),
    ("csg_coop_interval_rays", ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_int64)]),
    ("csg_camera_move", ctypes.c_int, [ctypes.POINTER(Camera), ctypes.c_float, ctypes.c_float, ctypes.c_float]),
    ("csg_camera_rotate", ctypes.c_int, [ctypes.POINTER(Camera), ctypes.c_float, ctypes.c_float]),
    ("csg_camera_input", ctypes.c_int, [ctypes.POINTER(Camera), ctypes.c_uint32, ctypes.c_int32, ctypes.c_float,
                                         ctypes.c_float]),
    ("csg_write_ppm", ctypes.c_int, [ctypes.c_char_p, _P, ctypes.c_int32, ctypes.c_int32]),
    ("csg_write_png", ctypes.c_int, [ctypes.c_char_p, _P, ctypes.c_int32, ctypes.c_int32]),
    ("csg_write_raw", ctypes.c_int, [ctypes.c_char_p, _P, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                     ctypes.c_int32]),
    ("csg_render_plan", ctypes.c_int, [_P, ctypes.POINTER(RenderOpts), ctypes.POINTER(LaunchInfo)]),
    ("csg_jitter", ctypes.c_float, [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32]),
    ("csg_render_batch", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.POINTER(RenderOpts), _P]),
    ("csg_batch_order", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_void_p]),
    ("csg_batch_create", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.POINTER(RenderOpts),
                                        ctypes.POINTER(_P)]),
    ("csg_batch_update", ctypes.c_int, [_P, ctypes.c_void_p, ctypes.c_int32, _P]),
    ("csg_batch_launch", ctypes.c_int, [_P, _P]),
    ("csg_batch_autotune", ctypes.c_int, [_P, ctypes.c_int32, _P, ctypes.POINTER(BatchTuning)]),
    ("csg_batch_free", ctypes.c_int, [_P]),
    ("csg_server_create", ctypes.c_int, [ctypes.POINTER(ServerConfig), ctypes.POINTER(_P)]),
    ("csg_server_add_scene", ctypes.c_int, [_P, _P, ctypes.POINTER(ctypes.c_int32)]),
    ("csg_server_submit", ctypes.c_int, [_P, ctypes.POINTER(Request), ctypes.POINTER(ctypes.c_int64)]),
    ("csg_server_poll", ctypes.c_int, [_P, ctypes.c_int64, ctypes.POINTER(ctypes.c_int32)]),
    ("csg_server_wait", ctypes.c_int, [_P, ctypes.c_int64, ctypes.POINTER(Outputs), ctypes.POINTER(RequestStats)]),
    ("csg_server_copy", ctypes.c_int, [_P, ctypes.c_int64, ctypes.POINTER(Outputs), _P]),
    ("csg_server_release", ctypes.c_int, [_P, ctypes.c_int64]),
    ("csg_server_stats_get", ctypes.c_int, [_P, ctypes.POINTER(ServerStats)]),
    ("csg_server_destroy", ctypes.c_int, [_P]),
    ("csg_last_error", ctypes.c_char_p, []),
    ("csg_version", ctypes.c_char_p, []),
]


def lib():
    """Load libcsgray.so (raises if it has not been built)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it first (python -c 'import __graft_entry__ as g; "
                               "g.build()' or make -C cuda-csg-tree-raycasting_amd)")
        L = ctypes.CDLL(LIB_PATH)
        for name, res, args in _SIGNATURES:
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = L
    return _LIB


def symbols():
    return [s[0] for s in _SIGNATURES]


def _check(status):
    if status != OK:
        raise CsgError(status, lib().csg_last_error().decode(errors="replace"))


def camera_default():
    c = Camera()
    lib().csg_camera_default(ctypes.byref(c))
    return c


def camera_look_at(eye, target, vfov_deg):
    c = Camera()
    _check(lib().csg_camera_look_at(c_float3(*eye), c_float3(*target), float(vfov_deg), ctypes.byref(c)))
    return c


def camera(pos, pitch, yaw, vfov):
    return Camera(c_float3(*pos), pitch, yaw, vfov)


def camera_basis(cam):
    out = (ctypes.c_float * 10)()
    _check(lib().csg_camera_basis(ctypes.byref(cam), out))
    return list(out)


def light_default():
    l = Light()
    lib().csg_light_default(ctypes.byref(l))
    return l


def light_from_angles(polar, azimuth):
    l = Light()
    lib().csg_light_from_angles(float(polar), float(azimuth), ctypes.byref(l))
    return l


def light(direction):
    return Light(c_float3(*direction))


def phong_default():
    p = Phong()
    lib().csg_phong_default(ctypes.byref(p))
    return p


def phong(ka, kd, ks, shininess, bg=(0.08, 0.08, 0.11, 1.0)):
    return Phong(ka, kd, ks, shininess, (ctypes.c_float * 4)(*bg))


def cam_tuple(cam):
    """(x, y, z, pitch, yaw, vfov) — the oracle's camera argument."""
    return (cam.pos[0], cam.pos[1], cam.pos[2], cam.pitch, cam.yaw, cam.vfov)


class Scene:
    """A parsed scene (csg_scene*).  upload=False keeps it host-only."""

    def __init__(self, text: str, upload: bool = True):
        data = text.encode() if isinstance(text, str) else bytes(text)
        h = _P()
        err = ctypes.create_string_buffer(512)
        fn = lib().csg_scene_parse if upload else lib().csg_scene_parse_host
        st = fn(data, len(data), ctypes.byref(h), err, len(err))
        if st != OK:
            raise CsgError(st, err.value.decode(errors="replace"))
        self.handle = h

    def upload(self):
        _check(lib().csg_scene_upload(self.handle))

    def close(self):
        if getattr(self, "handle", None):
            lib().csg_scene_free(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def info(self) -> SceneInfo:
        i = SceneInfo()
        _check(lib().csg_scene_info_get(self.handle, ctypes.byref(i)))
        return i

    def export(self):
        """(links int32[n,5], boxes float32[n,6], prims float32[p,14]) as numpy arrays."""
        import numpy as np
        i = self.info()
        links = np.zeros((i.nodes, 5), np.int32)
        boxes = np.zeros((i.nodes, 6), np.float32)
        prims = np.zeros((i.primitives, 14), np.float32)
        _check(lib().csg_scene_export(self.handle, links.ctypes.data, boxes.ctypes.data, prims.ctypes.data))
        return links, boxes, prims

    def plan(self, mode=MODE_FAST, block=0, placement=0, coop=0, coop_grace=0) -> LaunchInfo:
        info = LaunchInfo()
        opts = RenderOpts(mode, block, placement, 0, None, None, 0, 0, 0, coop, 0, coop_grace)
        _check(lib().csg_render_plan(self.handle, ctypes.byref(opts), ctypes.byref(info)))
        return info

    def render(self, cam, light_, phong_, width, height, mode=MODE_FAST, outputs=None, counters=False,
               stream=None, device=None, block=0, placement=0, tile_order=None, samples=1, seed=0, walk=0, coop=0,
               coop_when=0, coop_grace=0, coop_intervals=0, tight_boxes=0):
        """Render into torch CUDA tensors; returns dict rgba/depth/prim_id/flags(/counters).

        `outputs` may pass preallocated tensors (same keys) to avoid allocation.  `coop` is the
        step budget after which a ray may move to the cooperative kernel (0 default, < 0 never);
        `coop_when` 0 moves it only in the frame's tail (once every tile has been taken and a further
        `coop_grace` percent of the time that took has passed; 0 default, < 0 none), 1 as soon as the
        budget is spent.  `coop_intervals` < 0 keeps handed-over rays on the state machine (0: interval
        classification where the scene allows, fast mode); `tight_boxes` < 0 culls with the reference's
        boxes in fast mode too."""
        import torch
        dev = device or torch.device("cuda", torch.cuda.current_device())
        spp = max(1, int(samples))
        n = width * height * spp
        if outputs is None:
            outputs = alloc_outputs(width, height, spp, dev, counters)
        for k, t in outputs.items():
            assert t.is_cuda and t.is_contiguous(), k
        assert outputs["rgba"].numel() >= 4 * n if "rgba" in outputs else True
        o = Outputs(*(outputs[k].data_ptr() if outputs.get(k) is not None else None
                      for k in ("rgba", "depth", "prim_id", "flags")))
        opts = RenderOpts(mode, block, placement, 1 if counters else 0,
                          outputs["counters"].data_ptr() if counters else None,
                          tile_order.data_ptr() if tile_order is not None else None, spp, seed, walk, coop,
                          coop_when, coop_grace, coop_intervals, tight_boxes)
        s = stream if stream is not None else torch.cuda.current_stream(dev).cuda_stream
        _check(lib().csg_render_ex(self.handle, ctypes.byref(cam), ctypes.byref(light_), ctypes.byref(phong_),
                                   int(width), int(height), ctypes.byref(o), ctypes.byref(opts), s))
        return outputs


def alloc_outputs(width, height, spp=1, device=None, counters=False):
    """Output tensors for one render: (H, W*spp, ...) row-major, sample s of pixel x at column x*spp+s."""
    import torch
    dev = device or torch.device("cuda", torch.cuda.current_device())
    w = width * spp
    out = {
        "rgba": torch.empty((height, w, 4), dtype=torch.float32, device=dev),
        "depth": torch.empty((height, w), dtype=torch.float32, device=dev),
        "prim_id": torch.empty((height, w), dtype=torch.int32, device=dev),
        "flags": torch.empty((height, w), dtype=torch.uint8, device=dev),
    }
    if counters:
        out["counters"] = torch.zeros((height, w, 6), dtype=torch.int32, device=dev)
    return out


def _outputs_struct(outputs):
    return Outputs(*(outputs[k].data_ptr() if outputs.get(k) is not None else None
                     for k in ("rgba", "depth", "prim_id", "flags")))


KEY_FORWARD, KEY_BACKWARD, KEY_LEFT, KEY_RIGHT, KEY_UP, KEY_DOWN = 1, 2, 4, 8, 16, 32


def camera_move(cam, forward, right, up):
    """Camera::move in place (forward/right along the view, up along world Y)."""
    _check(lib().csg_camera_move(ctypes.byref(cam), forward, right, up))
    return cam


def camera_rotate(cam, dpitch, dyaw):
    """Camera::rotate in place (pitch clamped to +-89 degrees)."""
    _check(lib().csg_camera_rotate(ctypes.byref(cam), dpitch, dyaw))
    return cam


def camera_input(cam, keys=0, mouse=None):
    """One frame of the interactive input: KEY_* bits step 0.1; mouse=(dx, dy) rotates by -0.005/pixel."""
    dx, dy = mouse if mouse is not None else (0.0, 0.0)
    _check(lib().csg_camera_input(ctypes.byref(cam), int(keys), 1 if mouse is not None else 0, dx, dy))
    return cam


def _host(a, dtype):
    import numpy as np
    if hasattr(a, "detach"):
        a = a.detach().cpu().numpy()
    return np.ascontiguousarray(a, dtype=dtype)


def write_image(path, rgba):
    """Write an (H, W, 4) float RGBA render (row 0 = bottom) as .png or .ppm, top row first."""
    import numpy as np
    a = _host(rgba, np.float32)
    h, w = a.shape[0], a.shape[1]
    fn = lib().csg_write_png if str(path).lower().endswith(".png") else lib().csg_write_ppm
    _check(fn(str(path).encode(), a.ctypes.data, w, h))


def write_raw(path, arr, flip_rows=False):
    """Write an (H, W[, ...]) array (depth f32, prim id i32, flags u8, ...) as raw little-endian bytes."""
    import numpy as np
    a = np.ascontiguousarray(arr.detach().cpu().numpy() if hasattr(arr, "detach") else arr)
    h, w = a.shape[0], a.shape[1]
    _check(lib().csg_write_raw(str(path).encode(), a.ctypes.data, a.itemsize * (a.size // (h * w)), w, h,
                               1 if flip_rows else 0))


def coop_rays(stream=None):
    """Rays handed to the cooperative kernel so far on `stream` (waits for it)."""
    import torch
    s = stream if stream is not None else torch.cuda.current_stream().cuda_stream
    v = ctypes.c_int64()
    _check(lib().csg_coop_rays(s, ctypes.byref(v)))
    return v.value


def coop_interval_rays(stream=None):
    """Of the handed-over rays, those finished by interval classification (waits for the stream)."""
    import torch
    s = stream if stream is not None else torch.cuda.current_stream().cuda_stream
    v = ctypes.c_int64()
    _check(lib().csg_coop_interval_rays(s, ctypes.byref(v)))
    return v.value


def jitter(seed, pixel, sample, axis):
    return float(lib().csg_jitter(seed, pixel, sample, axis))


def make_jobs(specs):
    """specs: list of (Scene, cam, light, phong, width, height, outputs dict) -> ctypes array of Job."""
    arr = (Job * len(specs))()
    for i, (scene, cam, light_, phong_, w, h, outs) in enumerate(specs):
        arr[i] = Job(scene.handle, cam, light_, phong_, int(w), int(h), _outputs_struct(outs))
    return arr


def batch_order(jobs):
    order = (ctypes.c_int32 * len(jobs))()
    _check(lib().csg_batch_order(jobs, len(jobs), order))
    return list(order)


def render_batch(jobs, mode=MODE_FAST, block=0, walk=0, samples=1, seed=0, stream=None, coop=0, coop_when=0,
                 coop_grace=0, coop_intervals=0, tight_boxes=0):
    """One persistent launch over a ctypes Job array (see make_jobs)."""
    import torch
    opts = RenderOpts(mode, block, PLACE_GLOBAL, 0, None, None, samples, seed, walk, coop, coop_when, coop_grace,
                      coop_intervals, tight_boxes)
    s = stream if stream is not None else torch.cuda.current_stream().cuda_stream
    _check(lib().csg_render_batch(jobs, len(jobs), ctypes.byref(opts), s))


class Batch:
    """A captured batch (csg_batch*): jobs fixed at creation, graph replayed by launch()."""

    def __init__(self, jobs, mode=MODE_FAST, block=0, walk=0, samples=1, seed=0, coop=0, coop_when=0, coop_grace=0,
                 coop_intervals=0, tight_boxes=0):
        opts = RenderOpts(mode, block, PLACE_GLOBAL, 0, None, None, samples, seed, walk, coop, coop_when, coop_grace,
                          coop_intervals, tight_boxes)
        h = _P()
        _check(lib().csg_batch_create(jobs, len(jobs), ctypes.byref(opts), ctypes.byref(h)))
        self.handle = h
        self.jobs = jobs

    def _stream(self, stream):
        import torch
        return stream if stream is not None else torch.cuda.current_stream().cuda_stream

    def launch(self, stream=None):
        _check(lib().csg_batch_launch(self.handle, self._stream(stream)))

    def update(self, jobs, stream=None):
        _check(lib().csg_batch_update(self.handle, jobs, len(jobs), self._stream(stream)))
        self.jobs = jobs

    def autotune(self, reps=2, stream=None) -> BatchTuning:
        t = BatchTuning()
        _check(lib().csg_batch_autotune(self.handle, int(reps), self._stream(stream), ctypes.byref(t)))
        return t

    def close(self):
        if getattr(self, "handle", None):
            lib().csg_batch_free(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """csg_server*: resident scenes, request queue, continuous batching."""

    def __init__(self, budget_bytes=0, streams=0, max_batch_jobs=0, mode=MODE_FAST, coop=0, coop_when=0,
                 coop_grace=0):
        h = _P()
        cfg = ServerConfig(int(budget_bytes), int(streams), int(max_batch_jobs), int(mode), int(coop), int(coop_when),
                           int(coop_grace))
        _check(lib().csg_server_create(ctypes.byref(cfg), ctypes.byref(h)))
        self.handle = h
        self._scenes = []

    def add_scene(self, scene):
        i = ctypes.c_int32()
        _check(lib().csg_server_add_scene(self.handle, scene.handle, ctypes.byref(i)))
        self._scenes.append(scene)
        return i.value

    def submit(self, scene_id, cam, light_, phong_, width, height):
        t = ctypes.c_int64()
        r = Request(int(scene_id), cam, light_, phong_, int(width), int(height))
        _check(lib().csg_server_submit(self.handle, ctypes.byref(r), ctypes.byref(t)))
        return t.value

    def poll(self, ticket):
        d = ctypes.c_int32()
        _check(lib().csg_server_poll(self.handle, int(ticket), ctypes.byref(d)))
        return bool(d.value)

    def wait(self, ticket):
        st = RequestStats()
        _check(lib().csg_server_wait(self.handle, int(ticket), None, ctypes.byref(st)))
        return st

    def copy(self, ticket, outputs, stream=None):
        import torch
        s = stream if stream is not None else torch.cuda.current_stream().cuda_stream
        _check(lib().csg_server_copy(self.handle, int(ticket), ctypes.byref(_outputs_struct(outputs)), s))

    def release(self, ticket):
        _check(lib().csg_server_release(self.handle, int(ticket)))

    def stats(self):
        st = ServerStats()
        _check(lib().csg_server_stats_get(self.handle, ctypes.byref(st)))
        return st

    def close(self):
        if getattr(self, "handle", None):
            lib().csg_server_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


@dataclass
class RenderSpec:
    """One render job: scene text plus camera/light/shading and size."""
    text: str
    camera: Camera
    light: Light
    phong: Phong
    width: int
    height: int
