"""Serving loop on the GPU: every request's outputs are bit-identical to the
same request rendered alone, admission respects the memory budget, and
batches on different streams overlap on the device."""
import threading
import time

import numpy as np
import pytest
import torch

import csgray as C
from conftest import scene_text
from csgray import scenes as S

pytestmark = pytest.mark.gpu


def _resident(server, n=8, seed=21):
    mix = S.batch_mix(n, seed=seed)
    scenes = [C.Scene(t) for t, _, _ in mix]
    ids = [server.add_scene(s) for s in scenes]
    return scenes, ids


def _alone(scene, cam, light, ph, w, h):
    out = scene.render(cam, light, ph, w, h)
    torch.cuda.synchronize()
    return {k: v.cpu().numpy() for k, v in out.items()}


def test_results_equal_alone_and_in_the_mix(gpu):
    print("\n[case] 60 requests, 3 streams", flush=True)
    server = C.Server(streams=3, max_batch_jobs=16)
    scenes, ids = _resident(server)
    light, ph = C.light(S.LIGHT_TOWARD), C.phong(*S.PHONG)
    rng = np.random.default_rng(4)
    reqs = []
    for i in range(60):
        k = int(rng.integers(0, len(ids)))
        cam = C.camera_look_at(S.orbit_eye(rng, 4.0, 85.0), (0, 0, 0), 60)
        w, h = int(rng.integers(16, 260)), int(rng.integers(16, 200))
        reqs.append((k, cam, w, h, server.submit(ids[k], cam, light, ph, w, h)))
        if i % 7 == 0:
            time.sleep(0.002)  # arrivals while earlier batches run
    for k, cam, w, h, t in reqs:
        st = server.wait(t)
        assert st.status == C.OK and st.done_s >= st.start_s >= st.submit_s
        got = C.alloc_outputs(w, h)
        server.copy(t, got)
        torch.cuda.synchronize()
        server.release(t)
        ref = _alone(scenes[k], cam, light, ph, w, h)
        for key in ref:
            assert np.array_equal(ref[key].view(np.uint8), got[key].cpu().numpy().view(np.uint8)), key
    st = server.stats()
    assert st.completed == 60 and st.bytes_in_use == 0 and st.batches >= 2
    server.close()


def test_budget_admission_and_rejection(gpu):
    light, ph = C.light(S.LIGHT_TOWARD), C.phong(*S.PHONG)
    w = h = 256
    per = (16 + 4 + 4 + 1) * w * h
    budget = 2 * per + 4096
    server = C.Server(budget_bytes=budget, streams=2, max_batch_jobs=8)
    scene = C.Scene(scene_text("testCheese256.txt"))
    sid = server.add_scene(scene)
    with pytest.raises(C.CsgError) as e:
        server.submit(sid, C.camera_default(), light, ph, 1024, 1024)
    assert e.value.status == C.E_BUDGET
    tickets = [server.submit(sid, C.camera_default(), light, ph, w, h) for _ in range(12)]
    done = []

    def reaper():
        for t in tickets:
            st = server.wait(t)
            done.append(st.status)
            server.release(t)

    th = threading.Thread(target=reaper)
    th.start()
    th.join(timeout=60)
    assert not th.is_alive()
    assert done == [C.OK] * 12
    st = server.stats()
    assert st.peak_bytes_in_use <= budget and st.rejected == 1 and st.completed == 12
    server.close()


def test_batches_overlap_on_the_device(gpu):
    light, ph = C.light(S.LIGHT_TOWARD), C.phong(*S.PHONG)
    server = C.Server(streams=3, max_batch_jobs=4)
    scene = C.Scene(scene_text("testCheese512.txt"))
    sid = server.add_scene(scene)
    tickets = [server.submit(sid, C.camera_default(), light, ph, 512, 512) for _ in range(12)]
    stats = [server.wait(t) for t in tickets]
    for t in tickets:
        server.release(t)
    spans = sorted((s.start_s, s.done_s) for s in stats)
    overlap = any(b[0] < a[1] for a, b in zip(spans, spans[1:]) if b[0] != a[0])
    assert overlap, spans
    server.close()
