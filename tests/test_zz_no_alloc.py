"""csg_render_noalloc (the no-allocation variant of csg_render_ex, its scratch
in a static device array) against csg_render_ex, called through the C ABI:
two calls in a row, the second on a smaller frame, each bit-identical to
csg_render_ex with the same options (every output and the event counters);
the first hands every ray to the cooperative kernel, so the heavy-ray list
and resume area in the static array carry every ray of the frame.  Calls
above the array's 4,194,304 rays are refused before any launch."""
import ctypes

import numpy as np
import pytest
import torch

import csgray as C
from common import setups
from conftest import scene_text

pytestmark = pytest.mark.gpu


def _err():
    return C.lib().csg_last_error().decode(errors="replace")


def _call(fn, scene, cam, light, ph, w, h, mode, coop, when, spp=1):
    out = C.alloc_outputs(w, h, spp, counters=True)
    o = C.Outputs(*(out[k].data_ptr() for k in ("rgba", "depth", "prim_id", "flags")))
    opts = C.RenderOpts(mode, 0, 0, 1, out["counters"].data_ptr(), None, spp, 7, 0, coop, when, 0)
    st = fn(scene.handle, ctypes.byref(cam), ctypes.byref(light), ctypes.byref(ph), w, h, ctypes.byref(o),
            ctypes.byref(opts), torch.cuda.current_stream().cuda_stream)
    return st, out


@pytest.mark.parametrize("mode", [C.MODE_EXACT, C.MODE_FAST], ids=["exact", "fast"])
def test_noalloc_equals_render_ex_twice_second_smaller(gpu, mode):
    L = C.lib()
    sname, cam, light, ph, *_ = setups()[1]
    scene = C.Scene(scene_text("testCheese512.txt"))
    for w, h, coop, when, spp in ((640, 360, 1, 1, 1), (320, 180, 0, 0, 2)):
        print(f"\n[case] noalloc mode={mode} {w}x{h} spp={spp} coop={coop} when={when}", flush=True)
        st_a, a = _call(L.csg_render_noalloc, scene, cam, light, ph, w, h, mode, coop, when, spp)
        assert st_a == C.OK, _err()
        st_b, b = _call(L.csg_render_ex, scene, cam, light, ph, w, h, mode, coop, when, spp)
        assert st_b == C.OK, _err()
        torch.cuda.synchronize()
        assert int((b["prim_id"] >= 0).sum()) > 1000
        for k in ("rgba", "depth", "prim_id", "flags", "counters"):
            assert torch.equal(a[k].view(torch.uint8), b[k].view(torch.uint8)), k
    scene.close()


def test_noalloc_refuses_calls_above_its_scratch(gpu):
    L = C.lib()
    sname, cam, light, ph, *_ = setups()[1]
    scene = C.Scene(scene_text("testWikipedia.txt"))
    o = C.Outputs(None, None, None, None)
    opts = C.RenderOpts(C.MODE_FAST, 0, 0, 0, None, None, 1, 0, 0, 0, 0, 0)
    st = L.csg_render_noalloc(scene.handle, ctypes.byref(cam), ctypes.byref(light), ctypes.byref(ph), 2049, 2049,
                              ctypes.byref(o), ctypes.byref(opts), None)
    assert st == C.E_ARG and "rays per call" in _err()
    opts.samples = 3  # 1920 x 1080 x 3 = 6,220,800 rays
    st = L.csg_render_noalloc(scene.handle, ctypes.byref(cam), ctypes.byref(light), ctypes.byref(ph), 1920, 1080,
                              ctypes.byref(o), ctypes.byref(opts), None)
    assert st == C.E_ARG and "rays per call" in _err()
    scene.close()
