"""Serving loop on the GPU: every request's outputs are bit-identical to the
same request rendered alone and, at the serving configuration's request sizes
(256^2 to 1920x1080), to the FP32 oracle; admission respects the memory
budget, which counts the streams' workspaces; a block copied out and released
at once is not overwritten before the copy runs; and batches on different
streams overlap on the device."""
import threading
import time

import numpy as np
import pytest
import torch

import csgray as C
import oracle as O
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
    assert st.completed == 60 and st.bytes_in_use == st.workspace_bytes and st.batches >= 2
    server.close()


def _workspace_bytes(streams, max_batch_jobs):
    probe = C.Server(streams=streams, max_batch_jobs=max_batch_jobs)
    ws = probe.stats().workspace_bytes
    probe.close()
    return ws


def test_budget_admission_and_rejection(gpu):
    light, ph = C.light(S.LIGHT_TOWARD), C.phong(*S.PHONG)
    w = h = 256
    per = (16 + 4 + 4 + 1) * w * h
    ws = _workspace_bytes(2, 8)
    # each stream's queue, heavy-ray list with its resume area and job table
    assert ws >= 2 * 96 * 2**20
    with pytest.raises(C.CsgError) as e:
        C.Server(budget_bytes=ws, streams=2, max_batch_jobs=8)
    assert e.value.status == C.E_BUDGET
    budget = ws + 2 * per + 4096
    server = C.Server(budget_bytes=budget, streams=2, max_batch_jobs=8)
    assert server.stats().workspace_bytes == ws and server.stats().bytes_in_use == ws
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
    assert st.peak_bytes_in_use >= ws + per
    server.close()


def test_release_right_after_copy(gpu):
    """The client copies a result out on a busy stream of its own and releases
    the ticket at once, with no host synchronisation; the next request of the
    same size reuses the released block, and must not write it before that
    copy has run."""
    print("\n[case] copy then release, block reused", flush=True)
    light, ph = C.light(S.LIGHT_TOWARD), C.phong(*S.PHONG)
    server = C.Server(streams=1, max_batch_jobs=1)
    a_scene, b_scene = C.Scene(scene_text("testCheese256.txt")), C.Scene(scene_text("testWikipedia.txt"))
    a_id, b_id = server.add_scene(a_scene), server.add_scene(b_scene)
    w, h = 512, 512
    cam_a, cam_b = C.camera_default(), C.camera_look_at(S.VIEW_EYE, S.VIEW_TARGET, S.VIEW_VFOV_DEG)
    ta = server.submit(a_id, cam_a, light, ph, w, h)
    assert server.wait(ta).status == C.OK
    side = torch.cuda.Stream()
    dst = C.alloc_outputs(w, h)
    torch.cuda.synchronize()
    with torch.cuda.stream(side):
        torch.cuda._sleep(20_000_000)  # the copy waits behind this on the client's stream
    server.copy(ta, dst, stream=side.cuda_stream)
    server.release(ta)
    tb = server.submit(b_id, cam_b, light, ph, w, h)
    assert server.wait(tb).status == C.OK
    got_b = C.alloc_outputs(w, h)
    server.copy(tb, got_b)
    torch.cuda.synchronize()
    server.release(tb)
    ref_a = _alone(a_scene, cam_a, light, ph, w, h)
    ref_b = _alone(b_scene, cam_b, light, ph, w, h)
    for key in ref_a:
        assert np.array_equal(ref_a[key].view(np.uint8), dst[key].cpu().numpy().view(np.uint8)), key
        assert np.array_equal(ref_b[key].view(np.uint8), got_b[key].cpu().numpy().view(np.uint8)), key
    assert server.stats().batches == 2
    server.close()


def test_served_outputs_match_the_oracle_at_serving_sizes(gpu):
    """16 of the serving configuration's resident trees, requests of 256^2,
    512^2, 1024^2 and 1920x1080 with orbit cameras arriving about 1 ms apart,
    served in exact mode: every result equals the same request rendered alone,
    and on every 16th row the FP32 oracle (hit id, depth bits and flags
    exactly, RGBA within 2e-6)."""
    print("\n[case] 16 trees x serving sizes vs FP32 oracle", flush=True)
    light, ph = C.light(S.LIGHT_TOWARD), C.phong(*S.PHONG)
    sizes = [(256, 256), (512, 512), (1024, 1024), (1920, 1080)]
    mix = S.batch_mix(64, seed=64)[:16]
    server = C.Server(streams=3, max_batch_jobs=256, mode=C.MODE_EXACT)
    scenes = [C.Scene(t) for t, _, _ in mix]
    ids = [server.add_scene(sc) for sc in scenes]
    rng = np.random.default_rng(1000)
    reqs = []
    for i in range(16):
        w, h = sizes[i % 4]
        cam = C.camera_look_at(S.orbit_eye(rng, 4.0, 85.0), (0, 0, 0), 60)
        reqs.append((i, cam, w, h, server.submit(ids[i], cam, light, ph, w, h)))
        time.sleep(float(rng.exponential(1e-3)))
    served = []
    for i, cam, w, h, t in reqs:
        st = server.wait(t)
        assert st.status == C.OK
        got = C.alloc_outputs(w, h)
        server.copy(t, got)
        server.release(t)
        served.append(got)
    torch.cuda.synchronize()
    assert server.stats().batches >= 2
    for (i, cam, w, h, t), got in zip(reqs, served):
        print(f"[case] request {i}: tree {i}, {w}x{h}", flush=True)
        g = {k: v.cpu().numpy() for k, v in got.items()}
        alone = {k: v.cpu().numpy() for k, v in scenes[i].render(cam, light, ph, w, h, mode=C.MODE_EXACT).items()}
        torch.cuda.synchronize()
        for key in alone:
            assert np.array_equal(alone[key].view(np.uint8), g[key].view(np.uint8)), key
        rows = (0, h, 16)
        o = O.OracleScene(mix[i][0]).render(C.cam_tuple(cam), S.LIGHT_TOWARD, O.phong8(*S.PHONG), w, h, bits=32,
                                            rows=rows)
        sl = slice(0, h, 16)
        assert np.array_equal(g["prim_id"][sl], o["prim_id"][sl])
        assert np.array_equal(g["flags"][sl], o["flags"][sl])
        assert np.array_equal(g["depth"][sl].view(np.uint32), o["depth"][sl].view(np.uint32))
        assert np.abs(g["rgba"][sl] - o["rgba"][sl]).max() <= 2e-6
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
