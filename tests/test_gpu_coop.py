"""Cooperative evaluation of heavy rays on the GPU.

A ray whose state machine runs past the step budget is handed to a second
kernel that evaluates it with a whole wave (independent subtrees at once).
Every box test, primitive test and operator decision of the sequential
machine happens exactly once there, so the outputs and the per-ray event
counters must be bit-identical to the oracle and to the per-lane path,
whatever the budget.  With eager hand-over (coop_when=1) budget 1 hands
over every ray whose root is an operator and a budget of 16 splits a frame
between the two kernels; by default (coop_when=0) only rays still running
once every tile has been taken move, i.e. the frame's tail.
"""
import numpy as np
import pytest
import torch

import csgray as C
import oracle as O
from common import setups
from conftest import scene_names, scene_text
from csgray import scenes as S
import deep_trees as deep

pytestmark = pytest.mark.gpu


def _np(d):
    return {k: v.cpu().numpy() for k, v in d.items()}


def _same(a, b, keys=("rgba", "depth", "prim_id", "flags")):
    for k in keys:
        assert np.array_equal(a[k].view(np.uint8), b[k].view(np.uint8)), k


def _oracle_exact(text, cam6, l3, ph8, w, h, counters=True):
    return O.OracleScene(text).render(cam6, l3, ph8, w, h, bits=32, counters=counters)


@pytest.mark.parametrize("budget,when", [(1, 1), (16, 1), (1, 0)])
@pytest.mark.parametrize("setup", [0, 1])
@pytest.mark.parametrize("name", scene_names())
def test_deferred_rays_match_fp32_oracle_bitwise(gpu, name, setup, budget, when):
    print(f"\n[case] {name} setup={setup} budget={budget} when={when}", flush=True)
    sname, cam, light, phong, cam6, l3, ph8 = setups()[setup]
    w, h = 160, 90
    text = scene_text(name)
    s = C.Scene(text)
    g = _np(s.render(cam, light, phong, w, h, mode=C.MODE_EXACT, counters=True, coop=budget, coop_when=when))
    lane = _np(s.render(cam, light, phong, w, h, mode=C.MODE_EXACT, counters=True, coop=-1))
    torch.cuda.synchronize()
    s.close()
    o = _oracle_exact(text, cam6, l3, ph8, w, h)
    assert np.array_equal(g["prim_id"], o["prim_id"])
    assert np.array_equal(g["flags"], o["flags"])
    assert np.array_equal(g["depth"].view(np.uint32), o["depth"].view(np.uint32))
    assert np.array_equal(g["counters"], o["counters"])
    _same(g, lane, ("rgba", "depth", "prim_id", "flags", "counters"))


def test_cheese_split_between_kernels_matches_oracle(gpu):
    """The benchmark scene at 480x270 with the default budget and with 64
    (tail-only and eager): heavy rays go cooperative, the rest stay per lane."""
    print("\n[case] cheese 512 480x270 exact, budgets default/64, tail/eager", flush=True)
    sname, cam, light, phong, cam6, l3, ph8 = setups()[1]
    text = S.cheese(512, seed=512)
    s = C.Scene(text)
    o = _oracle_exact(text, cam6, l3, ph8, 480, 270)
    for budget, when in ((0, 0), (64, 0), (64, 1)):
        g = _np(s.render(cam, light, phong, 480, 270, mode=C.MODE_EXACT, counters=True, coop=budget, coop_when=when))
        heavy = int(((g["counters"][..., 0] + g["counters"][..., 1]) > (budget or s.plan(mode=C.MODE_EXACT).coop_budget)).sum())
        print(f"budget={budget} rays over budget={heavy}")
        assert heavy > 0
        assert np.array_equal(g["prim_id"], o["prim_id"])
        assert np.array_equal(g["flags"], o["flags"])
        assert np.array_equal(g["depth"].view(np.uint32), o["depth"].view(np.uint32))
        assert np.array_equal(g["counters"], o["counters"])
    s.close()


def test_fast_mode_same_with_and_without_deferral(gpu):
    """Fast mode, cheese 1920x1080: budgets 1, 64, default and never give the
    same bits in all four outputs."""
    print("\n[case] cheese 512 1920x1080 fast, budgets 1/64/default/never", flush=True)
    sname, cam, light, phong, cam6, l3, ph8 = setups()[1]
    s = C.Scene(S.cheese(512, seed=512))
    ref = _np(s.render(cam, light, phong, 1920, 1080, coop=-1))
    for budget, when in ((1, 1), (64, 1), (64, 0), (0, 0)):
        _same(_np(s.render(cam, light, phong, 1920, 1080, coop=budget, coop_when=when)), ref)
    s.close()


def test_more_handed_over_rays_than_direction_slots(gpu):
    """Fast mode, cheese 3840x2160, budget 1: more rays are handed over than
    the resume area keeps directions for (its last eighth, 16 B a ray: 512 Ki
    rays).  The first rays bring their direction to the cooperative kernel,
    the rest are rebuilt from their pixel; interval classification of both
    gives the per-lane kernel's bits."""
    print("\n[case] cheese 512 3840x2160 fast, budget 1 eager vs per-lane", flush=True)
    sname, cam, light, phong, cam6, l3, ph8 = setups()[1]
    s = C.Scene(S.cheese(512, seed=512))
    ref = _np(s.render(cam, light, phong, 3840, 2160, coop=-1))
    before, classified0 = C.coop_rays(), C.coop_interval_rays()
    got = _np(s.render(cam, light, phong, 3840, 2160, coop=1, coop_when=1))
    handed, classified = C.coop_rays() - before, C.coop_interval_rays() - classified0
    s.close()
    print(f"rays handed over: {handed}, classified by intervals: {classified}")
    assert handed > (64 << 20) // 8 // 16
    assert classified > 0.99 * handed
    _same(got, ref)


def _three_ways(s, cam, light, phong, w, h, **kw):
    """Per-lane only; every ray handed over (budget 1, eager) with interval
    classification; the same with the state machine.  Returns the three
    outputs and the rays the second render handed over and classified."""
    lane = _np(s.render(cam, light, phong, w, h, coop=-1, **kw))
    before, handed0 = C.coop_interval_rays(), C.coop_rays()
    iv = _np(s.render(cam, light, phong, w, h, coop=1, coop_when=1, **kw))
    classified, handed = C.coop_interval_rays() - before, C.coop_rays() - handed0
    machine = _np(s.render(cam, light, phong, w, h, coop=1, coop_when=1, coop_intervals=-1, **kw))
    return lane, iv, machine, classified, handed


@pytest.mark.parametrize("setup", [0, 1])
@pytest.mark.parametrize("name", scene_names())
def test_interval_classification_gives_the_state_machines_bits(gpu, name, setup):
    """Fast mode with every ray handed over: interval classification, the
    state-machine wave and the per-lane kernel give the same bits in all
    four outputs on every reference scene (coplanar faces and shared edges
    included: rays whose crossings tie go to the state machine), and scenes
    whose primitives all lie in their boxes are classified."""
    print(f"\n[case] {name} setup={setup} fast, intervals vs state machine", flush=True)
    sname, cam, light, phong, cam6, l3, ph8 = setups()[setup]
    s = C.Scene(scene_text(name))
    lane, iv, machine, classified, handed = _three_ways(s, cam, light, phong, 160, 90)
    inside = s.info().prims_in_boxes
    s.close()
    print(f"handed={handed} classified={classified} prims_in_boxes={inside}")
    _same(iv, lane)
    _same(machine, lane)
    if inside and handed > 0:
        assert classified > 0
    if not inside:
        assert classified == 0


@pytest.mark.parametrize("make", [lambda: S.cheese(512, seed=512), lambda: S.cheese(64, seed=3, r_lo=0.2, r_hi=0.5),
                                  lambda: deep.left_deep_difference(40, seed=2),
                                  lambda: deep.mixed_unbalanced(24, seed=5, cylinders=False)],
                         ids=["cheese512", "cheese64_big", "difference_chain40", "mixed24"])
def test_interval_classification_on_heavy_scenes(gpu, make):
    """The same three-way identity on scenes with long loops (the benchmark
    scene at 960x540 with 4 jittered samples, a coarse cheese, a 40-level
    Difference chain, a 24-level mixed tree); most handed-over rays are
    classified."""
    print("\n[case] heavy scene fast, intervals vs state machine, 960x540x4", flush=True)
    sname, cam, light, phong, cam6, l3, ph8 = setups()[1]
    s = C.Scene(make())
    lane, iv, machine, classified, handed = _three_ways(s, cam, light, phong, 960, 540, samples=4, seed=9)
    s.close()
    print(f"handed={handed} classified={classified}")
    _same(iv, lane)
    _same(machine, lane)
    assert handed > 0 and classified > 0.9 * handed


def _union_of_operators(n=96, seed=11):
    """A Union group whose members are operators: a balanced Union of n
    holed balls (Difference of two spheres), every fourth member an
    Intersection of a sphere and a cube, every eighth a Difference whose left
    operand is a Union of three spheres (a group of its own under it)."""
    rng = np.random.default_rng(seed)
    members = []
    for i in range(n):
        c = rng.uniform(-1.0, 1.0, 3)
        r = rng.uniform(0.15, 0.35)
        inner = c + rng.uniform(-0.1, 0.1, 3)
        if i % 8 == 7:
            balls = [S.sphere(c + rng.uniform(-0.15, 0.15, 3), "F5F500", r * 0.7) for _ in range(3)]
            members.append("\n".join(["Difference"] + S.balanced("Union", balls) + [S.sphere(inner, "F5F500", 0.5 * r)]))
        elif i % 4 == 3:
            members.append("\n".join(["Intersection", S.sphere(c, "E0E0E0", r), S.cube(c + 0.05, "E0E0E0", 1.5 * r)]))
        else:
            members.append("\n".join(["Difference", S.sphere(c, "F5F500", r), S.sphere(inner, "F5F500", 0.6 * r)]))
    return "\n".join(S.balanced("Union", members)) + "\n"


@pytest.mark.parametrize("make,groups,share", [(lambda: _union_of_operators(96, seed=11), 13, 0.9),
                                               (lambda: S.cheese(2048, seed=3), 2, 0.9)],
                         ids=["union_of_operators96", "cheese2048_two_groups"])
def test_interval_classification_across_union_groups(gpu, make, groups, share):
    """Union groups (flattened for the walk) whose members are operators
    sharing the group's slot, a group nested under a member, and a
    2048-sphere Union kept binary at the top over two 1024-member groups:
    classification, state machine and per-lane kernel agree bit for bit
    with every ray handed over.  (The 2048-sphere cheese is four times as
    dense as the benchmark's: about a third of its rays cross more than 31
    surfaces and are classified over stretches of the ray.)"""
    print("\n[case] union groups fast, intervals vs state machine, 480x270", flush=True)
    sname, cam, light, phong, cam6, l3, ph8 = setups()[1]
    s = C.Scene(make())
    assert s.info().union_groups == groups
    lane, iv, machine, classified, handed = _three_ways(s, cam, light, phong, 480, 270)
    s.close()
    print(f"handed={handed} classified={classified} hits={int((lane['prim_id'] >= 0).sum())}")
    _same(iv, lane)
    _same(machine, lane)
    assert handed > 0 and classified > share * handed


def test_batched_interval_classification_equals_per_lane(gpu):
    """Fast mode, a batch of scenes whose primitives lie in their boxes (cheese
    of 64, 512 and 2048 spheres -- the last with many rays of more than 31
    crossings, classified over stretches -- and a Union of holed balls), every
    ray handed over to the batched cooperative kernel: each job gets the bits
    of its own per-lane render, and the handed-over rays are classified."""
    print("\n[case] batch of interval-classified jobs, every ray handed over, fast", flush=True)
    light, ph = C.light(S.LIGHT_TOWARD), C.phong(*S.PHONG)
    sname, cam, light_, phong_, cam6, l3, ph8 = setups()[1]
    texts = [S.cheese(64, seed=1), S.cheese(512, seed=512), S.cheese(2048, seed=3), _union_of_operators(48, seed=5)]
    sizes = [(160, 120), (200, 150), (176, 99), (128, 128)]
    specs, alone = [], []
    for text, (w, h) in zip(texts, sizes):
        sc = C.Scene(text)
        alone.append(_np(sc.render(cam, light, ph, w, h, coop=-1)))
        specs.append((sc, cam, light, ph, w, h, C.alloc_outputs(w, h)))
    before, handed0 = C.coop_interval_rays(), C.coop_rays()
    C.render_batch(C.make_jobs(specs), coop=1, coop_when=1)
    torch.cuda.synchronize()
    classified, handed = C.coop_interval_rays() - before, C.coop_rays() - handed0
    print(f"handed={handed} classified={classified}")
    for spec, a in zip(specs, alone):
        _same(_np(spec[6]), a)
    assert handed > 0 and classified > 0.99 * handed
    for spec in specs:
        spec[0].close()


@pytest.mark.parametrize("eye,target", [((0.05, 0.02, 0.03), (1.0, 0.4, 0.2)), ((0.9, -0.7, 0.2), (-1.0, 0.3, -0.4))],
                         ids=["centre", "near_corner"])
def test_interval_classification_from_inside_the_solid(gpu, eye, target):
    """The camera inside the cheese cube: every ray starts inside the cube
    (its first crossing a lone Exit) and some inside spheres; classification,
    state machine and per-lane kernel agree bit for bit."""
    print(f"\n[case] cheese 512 from inside, eye={eye}, 480x270 fast", flush=True)
    light, phong = C.light(S.LIGHT_TOWARD), C.phong(*S.PHONG)
    s = C.Scene(S.cheese(512, seed=512))
    cam = C.camera_look_at(eye, target, 70)
    lane, iv, machine, classified, handed = _three_ways(s, cam, light, phong, 480, 270)
    s.close()
    print(f"handed={handed} classified={classified} hits={int((lane['prim_id'] >= 0).sum())}")
    _same(iv, lane)
    _same(machine, lane)
    assert handed > 0 and classified > 0.9 * handed


def test_heavy_list_overflow_finishes_per_lane(gpu):
    """2100x2100 rays with budget 1 overflow the 4 Mi-entry heavy list; the
    rays past its end finish in the per-lane kernel with the same results."""
    print("\n[case] Wikipedia 2100x2100 exact, budget 1 (heavy list overflow)", flush=True)
    sname, cam, light, phong, cam6, l3, ph8 = setups()[1]
    s = C.Scene(scene_text("testWikipedia.txt"))
    a = _np(s.render(cam, light, phong, 2100, 2100, mode=C.MODE_EXACT, counters=True, coop=1, coop_when=1))
    b = _np(s.render(cam, light, phong, 2100, 2100, mode=C.MODE_EXACT, counters=True, coop=-1))
    # a second call on the same stream starts from a clean heavy list
    c = _np(s.render(cam, light, phong, 2100, 2100, mode=C.MODE_EXACT, counters=True, coop=1, coop_when=1))
    s.close()
    _same(a, b, ("rgba", "depth", "prim_id", "flags", "counters"))
    _same(c, b, ("rgba", "depth", "prim_id", "flags", "counters"))


@pytest.mark.parametrize("target", [(0.99999, 0.0, 1.0), (0.99999, 0.99999, 1.0), (-0.99999, 0.99999, 1.0)],
                         ids=["edge", "corner", "corner_negative"])
def test_cube_normals_at_the_truncation_boundary(gpu, target):
    """The cube's face normal truncates p / half * 1.00001 (half = 1 here),
    which changes value at |p| = 0.99999: a camera with a 1e-4 degree field
    of view looking straight down at such a point puts every ray's hit within
    a few millionths of it, where the kernels take the IEEE division instead
    of the hardware reciprocal.  Exact mode equals the FP32 oracle bit for
    bit, every ray per lane and every ray handed over."""
    print(f"\n[case] cube face at {target}, 1e-4 degree view, exact", flush=True)
    text = "Union\nCube 0 0 0 FFFF00 2\nCube 4 4 4 00FFFF 0.1\n"
    eye = (target[0], target[1], 5.0)
    cam = C.camera_look_at(eye, target, 1e-4)
    sname, _, light, phong, _, l3, ph8 = setups()[1]
    s = C.Scene(text)
    o = O.OracleScene(text).render(C.cam_tuple(cam), l3, ph8, 64, 64, bits=32, counters=True)
    assert (o["prim_id"] == 0).all()
    for coop, when in ((-1, 0), (1, 1)):
        g = _np(s.render(cam, light, phong, 64, 64, mode=C.MODE_EXACT, counters=True, coop=coop, coop_when=when))
        assert np.array_equal(g["prim_id"], o["prim_id"])
        assert np.array_equal(g["flags"], o["flags"])
        assert np.array_equal(g["depth"].view(np.uint32), o["depth"].view(np.uint32))
        assert np.array_equal(g["counters"], o["counters"])
        assert np.abs(g["rgba"] - o["rgba"]).max() <= 2e-6
    fast = [_np(s.render(cam, light, phong, 64, 64, coop=coop, coop_when=when)) for coop, when in ((-1, 0), (1, 1))]
    s.close()
    _same(fast[0], fast[1])


@pytest.mark.parametrize("name", ["testWikipedia.txt", "testCheese512.txt"])
def test_deferred_jittered_samples_match_oracle(gpu, name):
    print(f"\n[case] {name} 4 spp budget 1", flush=True)
    sname, cam, light, phong, cam6, l3, ph8 = setups()[1]
    text = scene_text(name)
    s = C.Scene(text)
    w, h, spp, seed = 64, 48, 4, 1234
    g = _np(s.render(cam, light, phong, w, h, mode=C.MODE_EXACT, counters=True, samples=spp, seed=seed, coop=1,
                     coop_when=1))
    s.close()
    o = O.OracleScene(text).render_samples(cam6, l3, ph8, w, h, spp, seed, bits=32, counters=True)
    assert np.array_equal(g["prim_id"], o["prim_id"])
    assert np.array_equal(g["depth"].view(np.uint32), o["depth"].view(np.uint32))
    assert np.array_equal(g["flags"], o["flags"])
    assert np.array_equal(g["counters"], o["counters"])


def _batch_specs(n, seed, light, ph):
    rng = np.random.default_rng(seed)
    names = scene_names()
    out = []
    for i in range(n):
        text = scene_text(names[i % len(names)]) if i % 2 else S.random_tree(rng)
        eye = S.orbit_eye(rng, 5.0, 80.0)
        w, h = int(rng.integers(8, 160)), int(rng.integers(8, 120))
        sc = C.Scene(text)
        out.append((text, sc, C.camera_look_at(eye, (0, 0, 0), 60), w, h))
    return out


def test_batch_with_deferral_matches_oracle(gpu):
    print("\n[case] batch of 20 mixed jobs, exact, budget 1 and 16 eager, 1 tail", flush=True)
    light, ph = C.light(S.LIGHT_TOWARD), C.phong(*S.PHONG)
    specs = _batch_specs(20, 5, light, ph)
    for budget, when in ((1, 1), (16, 1), (1, 0)):
        full = [(sc, cam, light, ph, w, h, C.alloc_outputs(w, h)) for text, sc, cam, w, h in specs]
        C.render_batch(C.make_jobs(full), mode=C.MODE_EXACT, coop=budget, coop_when=when)
        torch.cuda.synchronize()
        for (text, sc, cam, w, h), f in zip(specs, full):
            g = _np(f[6])
            o = O.OracleScene(text).render(C.cam_tuple(cam), S.LIGHT_TOWARD, O.phong8(*S.PHONG), w, h, bits=32)
            assert np.array_equal(g["prim_id"], o["prim_id"])
            assert np.array_equal(g["depth"].view(np.uint32), o["depth"].view(np.uint32))
            assert np.array_equal(g["flags"], o["flags"])


def test_captured_batch_with_deferral_replays_and_updates(gpu):
    print("\n[case] captured batch, budget 1, replays and camera update", flush=True)
    light, ph = C.light(S.LIGHT_TOWARD), C.phong(*S.PHONG)
    specs = _batch_specs(12, 9, light, ph)
    full = [(sc, cam, light, ph, w, h, C.alloc_outputs(w, h)) for text, sc, cam, w, h in specs]
    want = [_np(sc.render(cam, light, ph, w, h, coop=-1)) for text, sc, cam, w, h in specs]
    b = C.Batch(C.make_jobs(full), coop=1, coop_when=1)
    for rep in range(3):
        for f in full:
            for t in f[6].values():
                t.fill_(0)
        torch.cuda.synchronize()
        b.launch()
        torch.cuda.synchronize()
        _ = torch.ones(1000, device="cuda").sum().item()  # eager work between replays
        for f, w_ in zip(full, want):
            _same(_np(f[6]), w_)
    # move every camera and resize; the graph's kernel nodes take the new jobs
    rng = np.random.default_rng(3)
    moved = []
    for sc, cam, l_, p_, w, h, outs in full:
        cam2 = C.camera_look_at(S.orbit_eye(rng, 4.0, 60.0), (0, 0, 0), 50)
        w2, h2 = int(rng.integers(8, 90)), int(rng.integers(8, 90))
        moved.append((sc, cam2, light, ph, w2, h2, C.alloc_outputs(w2, h2)))
    b.update(C.make_jobs(moved))
    b.launch()
    torch.cuda.synchronize()
    for sc, cam, l_, p_, w, h, outs in moved:
        _same(_np(outs), _np(sc.render(cam, light, ph, w, h, coop=-1)))
    b.close()


def test_plan_reports_cooperative_kernel(gpu):
    s = C.Scene(S.cheese(512, seed=512))
    p = s.plan()
    print(f"\ncoop budget={p.coop_budget} grid={p.coop_grid} lds={p.coop_lds_bytes}")
    # one wave's LDS: 16 B per operator (512: 511 unions and the difference) and a 6-byte entry (task,
    # tmin) per task-ring slot, as many as the operators without an operator child (256), 16-byte
    # rounded, then 16 finished rays waiting for their shading (16 B each)
    assert p.coop_budget > 0 and p.coop_grid >= p.compute_units
    assert p.coop_lds_bytes == (512 * 16 + 256 * 6 + 15) // 16 * 16 + 16 * 16
    # 16 one-wave groups per CU fit
    assert p.coop_grid == 16 * p.compute_units
    assert s.plan(coop=-1).coop_budget == 0
    assert s.plan(coop=7).coop_budget == 7
    s.close()


def test_default_tail_hand_over_engages_on_the_benchmark_frame(gpu):
    """With default options the cheese frame at 1920x1080 hands its tail to
    the cooperative kernel (tens of thousands of rays), while turning the
    hand-over off moves none; the images are identical either way."""
    print("\n[case] cheese 512 1920x1080 fast, default hand-over vs off", flush=True)
    sname, cam, light, phong, cam6, l3, ph8 = setups()[1]
    s = C.Scene(S.cheese(512, seed=512))
    before = C.coop_rays()
    a = _np(s.render(cam, light, phong, 1920, 1080))
    moved = C.coop_rays() - before
    b = _np(s.render(cam, light, phong, 1920, 1080, coop=-1))
    off = C.coop_rays() - before - moved
    s.close()
    print(f"rays handed over: {moved} (off: {off})")
    assert moved > 1000 and off == 0
    _same(a, b)


def test_every_bottom_operator_live_at_once(gpu):
    """1024 nearly concentric spheres under a balanced Union: a ray through the
    cluster hits every bounding box, so the 512 bottom operators (those without
    an operator child, which size the task ring) all have an ENTER live at
    once, far more than the wave's 64 lanes hold; every ray goes cooperative
    after one step and must match the oracle bit for bit."""
    print("\n[case] 1024 concentric spheres, exact, budget 1 eager", flush=True)
    rng = np.random.default_rng(77)
    c = rng.uniform(-0.01, 0.01, size=(1024, 3))
    r = rng.uniform(0.3, 1.0, size=1024)
    text = "\n".join(S.balanced("Union", [S.sphere(c[i], "F5F500", r[i]) for i in range(1024)])) + "\n"
    sname, cam, light, phong, cam6, l3, ph8 = setups()[1]
    s = C.Scene(text)
    p = s.plan()
    assert p.coop_lds_bytes == (1023 * 16 + 512 * 6 + 15) // 16 * 16 + 16 * 16
    w, h = 96, 54
    g = _np(s.render(cam, light, phong, w, h, mode=C.MODE_EXACT, counters=True, coop=1, coop_when=1))
    s.close()
    o = _oracle_exact(text, cam6, l3, ph8, w, h)
    # rays through the cluster test every operator's children
    assert int((o["counters"][..., 2] >= 2 * 1023).sum()) > 100
    assert np.array_equal(g["prim_id"], o["prim_id"])
    assert np.array_equal(g["flags"], o["flags"])
    assert np.array_equal(g["depth"].view(np.uint32), o["depth"].view(np.uint32))
    assert np.array_equal(g["counters"], o["counters"])
