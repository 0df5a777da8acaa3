"""ctypes binding of the oracle library (oracle/_build/libcsg_oracle.so).

ORACLE — test infrastructure only: the CPU restatement of the reference path
used to check the HIP product (tests/, __graft_entry__.smoke(), bench.py's
cpu_baseline leg).  The product never imports this module.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

ORACLE_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(ORACLE_DIR, "_build", "libcsg_oracle.so")
SAN_LIB_PATH = os.path.join(ORACLE_DIR, "_build", "libcsg_oracle_san.so")

# per-ray counter columns
COUNTER_NAMES = ("goto", "compute", "box", "sphere", "cylinder", "cube")
# degenerate-decision flag bits (combine tie, box/slab limit, silhouette or root at tmin, cylinder end,
# cube-normal truncation, tangent classification, stack overflow)
DG_TIE, DG_BOX, DG_ROOT, DG_CYL_END, DG_CUBE_N, DG_TANGENT, DG_OVERFLOW = 1, 2, 4, 8, 16, 32, 64

_LIB = None
_P = ctypes.c_void_p
_F = ctypes.POINTER(ctypes.c_float)


def build():
    subprocess.run(["make", "-s", "-C", ORACLE_DIR], check=True)


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        L.oracle_parse.restype = ctypes.c_int
        L.oracle_parse.argtypes = [ctypes.c_char_p, ctypes.c_size_t, ctypes.POINTER(_P), ctypes.c_char_p,
                                   ctypes.c_size_t]
        L.oracle_free.argtypes = [_P]
        L.oracle_info.argtypes = [_P, _P]
        L.oracle_dump_nodes.argtypes = [_P, _P, _P, _P]
        L.oracle_dump_prims.argtypes = [_P, _P, _P]
        L.oracle_camera_basis.argtypes = [_F, ctypes.c_int, _P]
        L.oracle_render.restype = ctypes.c_int
        L.oracle_render.argtypes = [_P, _F, _F, _F, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                    ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, _P, _P, _P, _P, _P, _P]
        L.oracle_classify.restype = ctypes.c_int
        L.oracle_classify.argtypes = [_P, _F, ctypes.c_int, ctypes.c_int, ctypes.c_double, _P, _P, _P, _P]
        L.oracle_set_mutation.argtypes = [ctypes.c_int]
        L.oracle_set_tangent_eta.argtypes = [ctypes.c_double]
        L.oracle_set_flag_margins.argtypes = [ctypes.c_double, ctypes.c_double]
        L.oracle_render_samples.restype = ctypes.c_int
        L.oracle_render_samples.argtypes = [_P, _F, _F, _F, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_uint32,
                                            ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                            _P, _P, _P, _P, _P, _P]
        L.oracle_jitter.restype = ctypes.c_float
        L.oracle_jitter.argtypes = [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32]
        _LIB = L
    return _LIB


def _f32(vals):
    return (ctypes.c_float * len(vals))(*[float(v) for v in vals])


class ParseError(ValueError):
    pass


class OracleScene:
    def __init__(self, text: str):
        data = text.encode() if isinstance(text, str) else bytes(text)
        h = _P()
        err = ctypes.create_string_buffer(512)
        if lib().oracle_parse(data, len(data), ctypes.byref(h), err, len(err)) != 0:
            raise ParseError(err.value.decode(errors="replace"))
        self.h = h

    def __del__(self):
        try:
            if self.h:
                lib().oracle_free(self.h)
        except Exception:
            pass

    def info(self):
        a = np.zeros(3, np.int32)
        lib().oracle_info(self.h, a.ctypes.data)
        return {"nodes": int(a[0]), "primitives": int(a[1]), "op_depth": int(a[2])}

    def nodes(self):
        n = self.info()["nodes"]
        links = np.zeros((n, 5), np.int32)
        b32 = np.zeros((n, 6), np.float32)
        b64 = np.zeros((n, 6), np.float64)
        lib().oracle_dump_nodes(self.h, links.ctypes.data, b32.ctypes.data, b64.ctypes.data)
        return links, b32, b64

    def prims(self):
        n = self.info()["primitives"]
        pf = np.zeros((n, 14), np.float32)
        ax = np.zeros((n, 3), np.float64)
        lib().oracle_dump_prims(self.h, pf.ctypes.data, ax.ctypes.data)
        return pf, ax

    def render(self, cam6, light3, phong8, width, height, bits=64, stack_cap=32, threads=0, rows=None,
               counters=False, degen=False):
        """Render with the restatement; returns dict of numpy arrays (H, W, ...).

        rows = (begin, end, step) restricts to a subset of image rows (others
        keep rgba 0 / depth nan / id -2 / flags 0)."""
        if threads <= 0:
            threads = min(os.cpu_count() or 1, 16)
        r0, r1, rs = rows if rows is not None else (0, height, 1)
        out = {
            "rgba": np.zeros((height, width, 4), np.float32),
            "depth": np.full((height, width), np.nan, np.float32),
            "prim_id": np.full((height, width), -2, np.int32),
            "flags": np.zeros((height, width), np.uint8),
        }
        if counters:
            out["counters"] = np.zeros((height, width, 6), np.int32)
        if degen:
            out["degen"] = np.zeros((height, width), np.uint8)
        st = lib().oracle_render(
            self.h, _f32(cam6), _f32(light3), _f32(phong8), width, height, bits, stack_cap, threads, r0, r1, rs,
            out["rgba"].ctypes.data, out["depth"].ctypes.data, out["prim_id"].ctypes.data, out["flags"].ctypes.data,
            out["counters"].ctypes.data if counters else None, out["degen"].ctypes.data if degen else None)
        if st != 0:
            raise ValueError("oracle_render: bad arguments")
        return out

    def render_samples(self, cam6, light3, phong8, width, height, spp, seed, bits=32, threads=0, rows=None,
                       counters=False, degen=False):
        """spp jittered rays per pixel; arrays shaped (H, W*spp, ...): sample s of pixel x at column x*spp+s."""
        if threads <= 0:
            threads = min(os.cpu_count() or 1, 16)
        r0, r1, rs = rows if rows is not None else (0, height, 1)
        w = width * spp
        out = {
            "rgba": np.zeros((height, w, 4), np.float32),
            "depth": np.full((height, w), np.nan, np.float32),
            "prim_id": np.full((height, w), -2, np.int32),
            "flags": np.zeros((height, w), np.uint8),
        }
        if counters:
            out["counters"] = np.zeros((height, w, 6), np.int32)
        if degen:
            out["degen"] = np.zeros((height, w), np.uint8)
        st = lib().oracle_render_samples(
            self.h, _f32(cam6), _f32(light3), _f32(phong8), width, height, spp, seed, bits, threads, r0, r1, rs,
            out["rgba"].ctypes.data, out["depth"].ctypes.data, out["prim_id"].ctypes.data, out["flags"].ctypes.data,
            out["counters"].ctypes.data if counters else None, out["degen"].ctypes.data if degen else None)
        if st != 0:
            raise ValueError("oracle_render_samples: bad arguments")
        return out

    def classify(self, cam6, width, height, sep=1e-6):
        st = np.zeros((height, width), np.int8)
        t = np.zeros((height, width), np.float64)
        pid = np.zeros((height, width), np.int32)
        en = np.zeros((height, width), np.uint8)
        if lib().oracle_classify(self.h, _f32(cam6), width, height, sep, st.ctypes.data, t.ctypes.data,
                                 pid.ctypes.data, en.ctypes.data) != 0:
            raise ValueError("oracle_classify: bad arguments")
        return {"status": st, "t": t, "prim_id": pid, "enter": en}


def jitter(seed, pixel, sample, axis):
    return float(lib().oracle_jitter(seed, pixel, sample, axis))


def set_tangent_eta(eta: float):
    """Near-tangent test of the flagger (for measurements; default 2e-6): the
    ray's closest approach within eta * distance of the radius; 0 = the
    discriminant within 1e-5 of b^2 + |c|."""
    lib().oracle_set_tangent_eta(float(eta))


def set_flag_margins(tie: float, box: float):
    """Relative margins of the flagger's combine-tie and box/slab tests (for
    measurements; defaults 1e-5 and 1e-6)."""
    lib().oracle_set_flag_margins(float(tie), float(box))


def set_mutation(m: int):
    """Switch on planted mistake `m` (0 = none); only for tests that show the pins catch it."""
    lib().oracle_set_mutation(int(m))


def camera_basis(cam6, bits=32):
    out = np.zeros(10, np.float64)
    lib().oracle_camera_basis(_f32(cam6), bits, out.ctypes.data)
    return out


def phong8(ka, kd, ks, n, bg=(0.08, 0.08, 0.11, 1.0)):
    return (ka, kd, ks, n) + tuple(bg)
