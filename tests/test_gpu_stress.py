"""The stress configuration on the GPU: cube minus a balanced Union of 4095
small spheres and cylinders (8191 nodes, 13 operator levels).  The scene does
not fit LDS, so the render reads it from global memory, and the cooperative
kernel holds 4095 operator states per wave (one wave per CU).  Exact mode
must match the FP32 oracle bit for bit (counters included) with jittered
samples, whether rays stay per lane or move to the cooperative kernel; the
plain launch and the captured batch of the same frame must agree.
"""
import numpy as np
import pytest
import torch

import csgray as C
import oracle as O
from csgray import scenes as S

pytestmark = pytest.mark.gpu


def _np(d):
    return {k: v.cpu().numpy() for k, v in d.items()}


@pytest.fixture(scope="module")
def stress():
    text = S.stress(4095, seed=4095)
    return text, C.Scene(text)


@pytest.mark.parametrize("placement", [C.PLACE_AUTO, C.PLACE_CACHE], ids=["auto", "lds-cache"])
@pytest.mark.parametrize("coop,when", [(-1, 0), (0, 0), (256, 1)])
def test_stress_samples_match_fp32_oracle(gpu, stress, coop, when, placement):
    print(f"\n[case] stress 8191 nodes 160x90 2 spp exact coop={coop} when={when} placement={placement}", flush=True)
    text, scene = stress
    cam = C.camera_look_at(S.VIEW_EYE, S.VIEW_TARGET, S.VIEW_VFOV_DEG)
    w, h, spp, seed = 160, 90, 2, 4095
    plan = scene.plan(mode=C.MODE_EXACT, coop=coop, placement=placement)
    assert not plan.scene_in_lds
    if placement == C.PLACE_CACHE:
        # part of the tree in LDS (level order: the top levels), the rest from global memory
        assert 0 < plan.cache_ops < scene.info().nodes // 2 and plan.lds_bytes <= plan.lds_limit
    g = _np(scene.render(cam, C.light(S.LIGHT_TOWARD), C.phong(*S.PHONG), w, h, mode=C.MODE_EXACT, counters=True,
                         samples=spp, seed=seed, coop=coop, coop_when=when, placement=placement))
    o = O.OracleScene(text).render_samples(C.cam_tuple(cam), S.LIGHT_TOWARD, O.phong8(*S.PHONG), w, h, spp, seed,
                                           bits=32, counters=True)
    assert np.array_equal(g["prim_id"], o["prim_id"])
    assert np.array_equal(g["depth"].view(np.uint32), o["depth"].view(np.uint32))
    assert np.array_equal(g["flags"], o["flags"])
    assert np.array_equal(g["counters"], o["counters"])
    assert np.abs(g["rgba"] - o["rgba"]).max() <= 2e-6


def test_stress_plain_equals_captured_batch(gpu, stress):
    print("\n[case] stress 640x360 4 spp fast: plain vs captured batch", flush=True)
    text, scene = stress
    cam = C.camera_look_at(S.VIEW_EYE, S.VIEW_TARGET, S.VIEW_VFOV_DEG)
    light, ph = C.light(S.LIGHT_TOWARD), C.phong(*S.PHONG)
    w, h, spp, seed = 640, 360, 4, 7
    plain = _np(scene.render(cam, light, ph, w, h, samples=spp, seed=seed))
    outs = C.alloc_outputs(w, h, spp)
    batch = C.Batch(C.make_jobs([(scene, cam, light, ph, w, h, outs)]), samples=spp, seed=seed)
    batch.autotune(reps=1)
    batch.launch()
    torch.cuda.synchronize()
    got = _np(outs)
    batch.close()
    for k in ("rgba", "depth", "prim_id", "flags"):
        assert np.array_equal(plain[k].view(np.uint8), got[k].view(np.uint8)), k
