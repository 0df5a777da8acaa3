"""Batched persistent launch and multi-sample rays on the GPU.

Every job of a batch must be bit-identical (all four outputs) to the same job
rendered alone, whatever else shares the launch; exact-mode jobs are also
bit-identical to the FP32 oracle (hit id, depth, flags, event counters).
Jittered multi-sample rays are checked against the oracle the same way.
"""
import numpy as np
import pytest
import torch

import csgray as C
import oracle as O
from common import setups
from conftest import scene_names, scene_text
from csgray import scenes as S

pytestmark = pytest.mark.gpu


def _np(d):
    return {k: v.cpu().numpy() for k, v in d.items()}


def mixed_jobs(n=24, seed=7):
    """Reference scenes and random trees with orbit cameras and mixed sizes."""
    rng = np.random.default_rng(seed)
    names = scene_names()
    texts = [scene_text(names[i % len(names)]) for i in range(n // 2)]
    texts += [S.random_tree(rng) for _ in range(n - n // 2)]
    out = []
    for t in texts:
        eye = S.orbit_eye(rng, 5.0, 80.0)
        w = int(rng.integers(2, 200))
        h = int(rng.integers(2, 120))
        out.append((t, eye, w, h))
    return out


def test_batch_jobs_equal_individual_renders_and_oracle(gpu):
    print("\n[case] batch of 24 mixed jobs, exact mode", flush=True)
    light, ph = C.light(S.LIGHT_TOWARD), C.phong(*S.PHONG)
    specs, alone, scenes = [], [], []
    for text, eye, w, h in mixed_jobs():
        sc = C.Scene(text)
        cam = C.camera_look_at(eye, (0, 0, 0), 60)
        scenes.append((sc, cam, text))
        alone.append(_np(sc.render(cam, light, ph, w, h, mode=C.MODE_EXACT, counters=True)))
        specs.append((sc, cam, light, ph, w, h, C.alloc_outputs(w, h, counters=True)))
    torch.cuda.synchronize()
    jobs = C.make_jobs(specs)
    opts = C.RenderOpts(C.MODE_EXACT, 0, C.PLACE_GLOBAL, 0, None, None, 1, 0, 0)
    # counters are per job; the batch API takes them through each job's own buffer only when
    # the opts buffer is set, so count through the plain path and compare the four outputs here
    C.render_batch(jobs, mode=C.MODE_EXACT)
    torch.cuda.synchronize()
    for (sc, cam, text), spec, a in zip(scenes, specs, alone):
        b = _np(spec[6])
        for k in ("rgba", "depth", "prim_id", "flags"):
            assert np.array_equal(a[k].view(np.uint8), b[k].view(np.uint8)), k
        w, h = spec[4], spec[5]
        o = O.OracleScene(text).render(C.cam_tuple(cam), S.LIGHT_TOWARD, O.phong8(*S.PHONG), w, h, bits=32,
                                       counters=True)
        assert np.array_equal(a["prim_id"], o["prim_id"])
        assert np.array_equal(a["depth"].view(np.uint32), o["depth"].view(np.uint32))
        assert np.array_equal(a["flags"], o["flags"])
        assert np.array_equal(a["counters"], o["counters"])


def test_batch_fast_mode_equals_individual_fast_renders(gpu):
    light, ph = C.light(S.LIGHT_TOWARD), C.phong(*S.PHONG)
    specs, alone = [], []
    for text, eye, w, h in mixed_jobs(40, seed=11):
        sc = C.Scene(text)
        cam = C.camera_look_at(eye, (0, 0, 0), 60)
        alone.append(_np(sc.render(cam, light, ph, w, h)))
        specs.append((sc, cam, light, ph, w, h, C.alloc_outputs(w, h)))
    jobs = C.make_jobs(specs)
    order = C.batch_order(jobs)
    jobs_sorted = C.make_jobs([specs[i] for i in order])
    C.render_batch(jobs_sorted, walk=1)
    torch.cuda.synchronize()
    for spec, a in zip(specs, alone):
        b = _np(spec[6])
        for k in a:
            assert np.array_equal(a[k].view(np.uint8), b[k].view(np.uint8)), k


def test_batch_order_is_heaviest_first(gpu):
    light, ph = C.light(S.LIGHT_TOWARD), C.phong(*S.PHONG)
    scs = [C.Scene(scene_text(n)) for n in ("test1.txt", "testCheese512.txt", "test5.txt")]
    sizes = [(64, 64), (64, 64), (256, 256)]
    specs = [(sc, C.camera_default(), light, ph, w, h, C.alloc_outputs(w, h)) for sc, (w, h) in zip(scs, sizes)]
    assert C.batch_order(C.make_jobs(specs)) == [2, 1, 0]


def test_more_jobs_than_one_launch_takes(gpu):
    light, ph = C.light(S.LIGHT_TOWARD), C.phong(*S.PHONG)
    sc = C.Scene(scene_text("testWikipedia.txt"))
    rng = np.random.default_rng(3)
    specs = []
    for i in range(1100):
        cam = C.camera_look_at(S.orbit_eye(rng, 5.0, 80.0), (0, 0, 0), 60)
        specs.append((sc, cam, light, ph, 9, 7, C.alloc_outputs(9, 7)))
    C.render_batch(C.make_jobs(specs))
    torch.cuda.synchronize()
    for i in (0, 1023, 1024, 1099):
        sc_, cam, l_, p_, w, h, outs = specs[i]
        a = _np(sc.render(cam, light, ph, w, h))
        torch.cuda.synchronize()
        b = _np(outs)
        assert np.array_equal(a["prim_id"], b["prim_id"]) and np.array_equal(a["depth"], b["depth"])


@pytest.mark.parametrize("name", ["testWikipedia.txt", "testCylinderSphere.txt", "testCheese256.txt",
                                  "testSingleCube.txt"])
def test_jittered_samples_match_oracle(gpu, name):
    print(f"\n[case] {name} spp=4", flush=True)
    sname, cam, light, phong, cam6, l3, ph8 = setups()[1]
    text = scene_text(name)
    w, h, spp, seed = 96, 54, 4, 1234
    s = C.Scene(text)
    g = _np(s.render(cam, light, phong, w, h, mode=C.MODE_EXACT, samples=spp, seed=seed, counters=True))
    torch.cuda.synchronize()
    o = O.OracleScene(text).render_samples(cam6, l3, ph8, w, h, spp, seed, bits=32, counters=True)
    assert np.array_equal(g["prim_id"], o["prim_id"])
    assert np.array_equal(g["depth"].view(np.uint32), o["depth"].view(np.uint32))
    assert np.array_equal(g["flags"], o["flags"])
    assert np.array_equal(g["counters"], o["counters"])
    assert np.abs(g["rgba"] - o["rgba"]).max() <= 2e-6


def test_morton_walk_changes_nothing(gpu):
    sname, cam, light, phong, *_ = setups()[1]
    s = C.Scene(scene_text("testCheese512.txt"))
    for (w, h) in ((256, 256), (200, 130)):
        a = _np(s.render(cam, light, phong, w, h, walk=0))
        b = _np(s.render(cam, light, phong, w, h, walk=1))
        torch.cuda.synchronize()
        for k in a:
            assert np.array_equal(a[k].view(np.uint8), b[k].view(np.uint8)), k


def _specs_for(mix, light, ph, size_override=None):
    specs = []
    for text, eye, size in mix:
        sc = C.Scene(text)
        cam = C.camera_look_at(eye, (0, 0, 0), 60)
        w = h = size_override or size
        specs.append((sc, cam, light, ph, w, h, C.alloc_outputs(w, h)))
    return specs


def _reference(specs):
    ref = []
    for sc, cam, l, p, w, h, _ in specs:
        ref.append(_np(sc.render(cam, l, p, w, h)))
    torch.cuda.synchronize()
    return ref


def _check(specs, ref):
    for spec, a in zip(specs, ref):
        b = _np(spec[6])
        for k in a:
            assert np.array_equal(a[k].view(np.uint8), b[k].view(np.uint8)), k


def test_captured_batch_replays_match_eager(gpu):
    print("\n[case] captured batch of 16, 3 replays", flush=True)
    light, ph = C.light(S.LIGHT_TOWARD), C.phong(*S.PHONG)
    mix = S.batch_mix(16, seed=99, resolutions=(32, 64, 96))
    specs = _specs_for(mix, light, ph)
    ref = _reference(specs)
    batch = C.Batch(C.make_jobs(specs))
    for rep in range(3):
        for spec in specs:
            for t in spec[6].values():
                t.fill_(0)
        torch.cuda.synchronize()
        batch.launch()
        torch.cuda.synchronize()
        _ = torch.ones(1000, device="cuda").sum().item()  # eager work between replays
        _check(specs, ref)


def test_captured_batch_updates_cameras_and_sizes(gpu):
    light, ph = C.light(S.LIGHT_TOWARD), C.phong(*S.PHONG)
    mix = S.batch_mix(12, seed=5, resolutions=(40, 72))
    specs = _specs_for(mix, light, ph)
    batch = C.Batch(C.make_jobs(specs))
    rng = np.random.default_rng(8)
    for rep in range(3):
        new = []
        for sc, cam, l, p, w, h, _ in specs:
            cam2 = C.camera_look_at(S.orbit_eye(rng, 4.0, 80.0), (0, 0, 0), 60)
            w2, h2 = (w, h) if rep == 0 else (int(rng.integers(8, 90)), int(rng.integers(8, 90)))
            new.append((sc, cam2, l, p, w2, h2, C.alloc_outputs(w2, h2)))
        ref = _reference(new)
        batch.update(C.make_jobs(new))
        batch.launch()
        torch.cuda.synchronize()
        _check(new, ref)


def test_captured_batch_updates_to_other_scenes_and_job_counts(gpu):
    """Updating a captured batch to different trees (deeper, with more
    operators: a larger stack slab and cooperative state, so the graph is
    captured again), then to fewer jobs, then back: every launch equals the
    eager renders of the same jobs, and exact mode equals the oracle."""
    print("\n[case] captured batch updated across scene sets and job counts", flush=True)
    light, ph = C.light(S.LIGHT_TOWARD), C.phong(*S.PHONG)
    shallow = _specs_for(S.batch_mix(10, seed=21, resolutions=(32, 64)), light, ph)
    rng = np.random.default_rng(22)
    deep_texts = [S.cheese(int(n), seed=int(n)) for n in rng.integers(200, 600, size=6)]
    deep = [(C.Scene(t), C.camera_look_at(S.orbit_eye(rng, 4.0, 60.0), (0, 0, 0), 60), light, ph, 48, 40,
             C.alloc_outputs(48, 40)) for t in deep_texts]
    batch = C.Batch(C.make_jobs(shallow))
    for specs in (shallow, deep, deep[:2], shallow[:7], deep):
        for spec in specs:
            for t in spec[6].values():
                t.fill_(0)
        ref = _reference(specs)
        batch.update(C.make_jobs(specs))
        batch.launch()
        torch.cuda.synchronize()
        _check(specs, ref)
    batch.close()
    exact = C.Batch(C.make_jobs(deep[:3]), mode=C.MODE_EXACT)
    exact.launch()
    torch.cuda.synchronize()
    for (sc, cam, l, p, w, h, outs), text in zip(deep[:3], deep_texts[:3]):
        g = _np(outs)
        o = O.OracleScene(text).render(C.cam_tuple(cam), S.LIGHT_TOWARD, O.phong8(*S.PHONG), w, h, bits=32)
        assert np.array_equal(g["prim_id"], o["prim_id"])
        assert np.array_equal(g["depth"].view(np.uint32), o["depth"].view(np.uint32))
        assert np.array_equal(g["flags"], o["flags"])
    exact.close()


def test_autotune_keeps_results(gpu):
    light, ph = C.light(S.LIGHT_TOWARD), C.phong(*S.PHONG)
    mix = S.batch_mix(20, seed=17, resolutions=(64, 128))
    specs = _specs_for(mix, light, ph)
    ref = _reference(specs)
    batch = C.Batch(C.make_jobs(specs))
    t = batch.autotune(reps=1)
    print(f"\nchosen block={t.block} walk={t.tile_walk} heaviest_first={t.heaviest_first} ms={t.ms:.3f}")
    assert t.block in (256, 512, 1024) and t.tile_walk in (0, 1) and t.ms > 0
    for spec in specs:
        for x in spec[6].values():
            x.fill_(0)
    batch.launch()
    torch.cuda.synchronize()
    _check(specs, ref)


def test_autotune_depth_classes(gpu):
    """A batch of shallow random trees and deep operator chains (32 levels):
    the autotuner also times the batch split into two depth classes, each
    launch's LDS stacks sized to its own deepest tree.  The split
    configuration is timed last, so the outputs right after tuning come from
    it; they, and the outputs of the chosen configuration, equal the solo
    renders, and in exact mode the oracle."""
    from deep_trees import left_deep_difference, right_deep_union
    print("\n[case] autotune with depth classes (shallow trees + 32-level chains)", flush=True)
    light, ph = C.light(S.LIGHT_TOWARD), C.phong(*S.PHONG)
    texts = [t for t, _, _ in S.batch_mix(10, seed=31)]
    texts += [left_deep_difference(32, seed=s) for s in range(3)] + [right_deep_union(32, seed=s) for s in range(3)]
    rng = np.random.default_rng(32)
    specs = []
    for t in texts:
        cam = C.camera_look_at(S.orbit_eye(rng, 4.5, 70.0), (0, 0, 0), 60)
        w, h = int(rng.integers(48, 160)), int(rng.integers(48, 120))
        specs.append((C.Scene(t), cam, light, ph, w, h, C.alloc_outputs(w, h)))
    depths = sorted(sc.info().op_depth for sc, *_ in specs)
    assert depths[0] <= 10 and depths[-1] == 32
    ref = [_np(sc.render(cam, l, p, w, h, mode=C.MODE_EXACT)) for sc, cam, l, p, w, h, _ in specs]
    torch.cuda.synchronize()
    batch = C.Batch(C.make_jobs(specs), mode=C.MODE_EXACT)
    t = batch.autotune(reps=1)
    torch.cuda.synchronize()
    print(f"chosen block={t.block} walk={t.tile_walk} heaviest_first={t.heaviest_first} "
          f"depth_split={t.depth_split} ms={t.ms:.3f}")
    assert t.depth_split in (0,) + tuple(d for d in depths if d < 32)
    _check(specs, ref)  # the last timed configuration: two depth classes
    for spec in specs:
        for x in spec[6].values():
            x.fill_(0)
    batch.launch()
    torch.cuda.synchronize()
    _check(specs, ref)
    batch.close()
    for (sc, cam, l, p, w, h, outs), text in zip(specs, texts):
        g = _np(outs)
        o = O.OracleScene(text).render(C.cam_tuple(cam), S.LIGHT_TOWARD, O.phong8(*S.PHONG), w, h, bits=32,
                                       stack_cap=2 * sc.info().op_depth + 1)
        assert np.array_equal(g["prim_id"], o["prim_id"])
        assert np.array_equal(g["depth"].view(np.uint32), o["depth"].view(np.uint32))
