"""GPU parity: the HIP render (through the C ABI) against the oracle.

Exact mode must be bit-identical to the oracle's FP32 build (hit id, depth
bits, flags, and the per-ray event counters of the state machine); RGBA may
differ by the ULPs of the device powf/sqrtf vs host libm.  Fast mode is held
to the FP64 oracle: hit id and enter/exit/miss agree on >= 99.99% of the rays
the FP64 run does not flag as degenerate, and depth is within 1e-4 relative.
"""
import numpy as np
import pytest
import torch

import csgray as C
import oracle as O
from common import setups
from conftest import scene_names, scene_text

pytestmark = pytest.mark.gpu

RGBA_TOL = 2e-6        # device vs host powf in the specular term
DEPTH_REL_TOL = 1e-4   # fast mode vs FP64
AGREE_MIN = 0.9999     # hit-id agreement on unflagged rays


def gpu_render(text, cam, light, phong, w, h, mode, counters=False, placement=C.PLACE_AUTO):
    s = C.Scene(text)
    out = s.render(cam, light, phong, w, h, mode=mode, counters=counters, placement=placement)
    torch.cuda.synchronize()
    res = {k: v.cpu().numpy() for k, v in out.items()}
    s.close()
    return res


def _label(name, setup):
    print(f"\n[case] {name} setup={setup}", flush=True)


@pytest.mark.parametrize("placement", [C.PLACE_GLOBAL, C.PLACE_LDS, C.PLACE_CACHE])
@pytest.mark.parametrize("setup", [0, 1])
@pytest.mark.parametrize("name", scene_names())
def test_exact_matches_fp32_oracle_bitwise(gpu, name, setup, placement):
    _label(name, f"{setup} placement={placement}")
    sname, cam, light, phong, cam6, l3, ph8 = setups()[setup]
    w, h = 160, 90
    text = scene_text(name)
    g = gpu_render(text, cam, light, phong, w, h, C.MODE_EXACT, counters=True, placement=placement)
    o = O.OracleScene(text).render(cam6, l3, ph8, w, h, bits=32, counters=True)
    assert np.array_equal(g["prim_id"], o["prim_id"])
    assert np.array_equal(g["flags"], o["flags"])
    assert np.array_equal(g["depth"].view(np.uint32), o["depth"].view(np.uint32))
    assert np.array_equal(g["counters"], o["counters"])
    assert np.abs(g["rgba"] - o["rgba"]).max() <= RGBA_TOL


def compare_fp64(g, o64, rows=None):
    """(hit-id + enter/exit/miss agreement on unflagged rays, max relative depth
    error on unflagged agreeing hits, flagged fraction).  `rows` restricts the
    comparison to the rows the oracle rendered."""
    sel = np.ones(o64["prim_id"].shape, bool) if rows is None else np.zeros(o64["prim_id"].shape, bool)
    if rows is not None:
        sel[rows[0]:rows[1]:rows[2]] = True
    unflagged = (o64["degen"] == 0) & sel
    cls = C.ENTER | C.EXIT | C.MISS
    agree = (g["prim_id"] == o64["prim_id"]) & ((g["flags"] & cls) == (o64["flags"] & cls))
    n_unf = int(unflagged.sum())
    rate = float((agree & unflagged).sum()) / max(n_unf, 1)
    hit = agree & unflagged & (o64["prim_id"] >= 0)
    rel = np.abs(g["depth"][hit].astype(np.float64) - o64["depth"][hit]) / np.abs(o64["depth"][hit])
    return rate, (float(rel.max()) if rel.size else 0.0), float(1 - unflagged.sum() / max(sel.sum(), 1))


def accuracy_report(label, g, o64, rows=None):
    """Prints the per-configuration accuracy figures: hit-id agreement on all
    and on unflagged rays, inside/outside (hit vs miss) agreement, flagged
    fraction, and depth relative error max / p99.9 on agreeing hits."""
    sel = np.ones(o64["prim_id"].shape, bool)
    if rows is not None:
        sel = np.zeros(o64["prim_id"].shape, bool)
        sel[rows[0]:rows[1]:rows[2]] = True
    unf = (o64["degen"] == 0) & sel
    same_id = g["prim_id"] == o64["prim_id"]
    same_hit = (g["prim_id"] >= 0) == (o64["prim_id"] >= 0)
    hit = same_id & sel & (o64["prim_id"] >= 0)
    rel = np.abs(g["depth"][hit].astype(np.float64) - o64["depth"][hit]) / np.abs(o64["depth"][hit])
    n = int(sel.sum())
    print(f"[accuracy] {label}: rays={n} id_agree_all={same_id[sel].mean():.6f} "
          f"id_agree_unflagged={same_id[unf].mean():.6f} inside_outside_agree={same_hit[sel].mean():.6f} "
          f"flagged={1 - unf.sum() / n:.4%} depth_rel_max={rel.max() if rel.size else 0:.2e} "
          f"depth_rel_p999={np.quantile(rel, 0.999) if rel.size else 0:.2e}", flush=True)


def test_fast_vs_fp64_wikipedia_full_size(gpu):
    """The Wikipedia configuration (640x360 from (3, 2.5, 4)) in fast mode against FP64."""
    _label("testWikipedia.txt 640x360 fast", 1)
    sname, cam, light, phong, cam6, l3, ph8 = setups()[1]
    text = scene_text("testWikipedia.txt")
    g = gpu_render(text, cam, light, phong, 640, 360, C.MODE_FAST)
    o64 = O.OracleScene(text).render(cam6, l3, ph8, 640, 360, bits=64, degen=True)
    accuracy_report("wikipedia 640x360 fast", g, o64)
    rate, drel, flagged = compare_fp64(g, o64)
    assert rate >= AGREE_MIN and drel <= DEPTH_REL_TOL


@pytest.mark.parametrize("setup", [0, 1])
@pytest.mark.parametrize("name", scene_names())
def test_fast_vs_fp64_oracle(gpu, name, setup):
    _label(name, setup)
    sname, cam, light, phong, cam6, l3, ph8 = setups()[setup]
    w, h = 320, 180
    text = scene_text(name)
    g = gpu_render(text, cam, light, phong, w, h, C.MODE_FAST)
    o64 = O.OracleScene(text).render(cam6, l3, ph8, w, h, bits=64, degen=True)
    rate, drel, flagged = compare_fp64(g, o64)
    print(f"agree(unflagged)={rate:.6f} depth_rel_max={drel:.2e} flagged={flagged:.4%}")
    assert rate >= AGREE_MIN
    assert drel <= DEPTH_REL_TOL


def test_exact_vs_fp64_wikipedia_full_size(gpu):
    """The Wikipedia scene at 640x360 from (3, 2.5, 4): exact mode against FP64."""
    _label("testWikipedia.txt 640x360", 1)
    sname, cam, light, phong, cam6, l3, ph8 = setups()[1]
    text = scene_text("testWikipedia.txt")
    g = gpu_render(text, cam, light, phong, 640, 360, C.MODE_EXACT)
    o32 = O.OracleScene(text).render(cam6, l3, ph8, 640, 360, bits=32)
    assert np.array_equal(g["prim_id"], o32["prim_id"])
    o64 = O.OracleScene(text).render(cam6, l3, ph8, 640, 360, bits=64, degen=True)
    accuracy_report("wikipedia 640x360 exact", g, o64)
    rate, drel, flagged = compare_fp64(g, o64)
    print(f"agree(unflagged)={rate:.6f} depth_rel_max={drel:.2e} flagged={flagged:.4%}")
    assert rate >= AGREE_MIN and drel <= DEPTH_REL_TOL


def test_fast_vs_fp64_cheese512_spheres_1080p(gpu):
    """The benchmark frame (cheese of 512 spheres, 1920x1080, from (3, 2.5, 4))
    in fast mode against the FP64 oracle on every ray."""
    from csgray import scenes as S
    _label("cheese 512 1920x1080", 1)
    sname, cam, light, phong, cam6, l3, ph8 = setups()[1]
    text = S.cheese(512, seed=512)
    g = gpu_render(text, cam, light, phong, 1920, 1080, C.MODE_FAST)
    o64 = O.OracleScene(text).render(cam6, l3, ph8, 1920, 1080, bits=64, degen=True)
    accuracy_report("cheese 1920x1080 fast (all rows)", g, o64)
    rate, drel, flagged = compare_fp64(g, o64)
    print(f"agree(unflagged)={rate:.6f} depth_rel_max={drel:.2e} flagged={flagged:.4%}")
    assert rate >= AGREE_MIN and drel <= DEPTH_REL_TOL
    assert not (g["flags"] & 128).any()


def _sampled(d, rows, keys=("prim_id", "flags", "depth", "degen")):
    """The rendered rows r0:r1:rs of a frame's arrays, flattened."""
    return {k: d[k][rows[0]:rows[1]:rows[2]].reshape(-1) for k in keys if k in d}


def _stacked(parts):
    return {k: np.concatenate([p[k] for p in parts]) for k in parts[0]}


def test_fast_vs_fp64_batch_mix(gpu):
    """The 256 random trees of the batched benchmark (128^2 to 1024^2, orbit
    cameras) rendered as one batched launch in fast mode, against the FP64
    oracle on every 16th row of every job (5.8 M rays)."""
    from csgray import scenes as S
    _label("256-tree batch mix", 1)
    light, ph = C.light(S.LIGHT_TOWARD), C.phong(*S.PHONG)
    mix = S.batch_mix(256, seed=256)
    specs = [(C.Scene(text), C.camera_look_at(eye, (0, 0, 0), 60), light, ph, size, size, C.alloc_outputs(size, size))
             for text, eye, size in mix]
    C.render_batch(C.make_jobs(specs))
    torch.cuda.synchronize()
    gs, os_ = [], []
    for (text, eye, size), spec in zip(mix, specs):
        rows = (0, size, 16)
        g = {k: spec[6][k][rows[0]:rows[1]:rows[2]].cpu().numpy().reshape(-1) for k in ("prim_id", "flags", "depth")}
        o64 = O.OracleScene(text).render(C.cam_tuple(spec[1]), S.LIGHT_TOWARD, O.phong8(*S.PHONG), size, size,
                                          bits=64, degen=True, rows=rows)
        gs.append(g)
        os_.append(_sampled(o64, rows))
    g, o64 = _stacked(gs), _stacked(os_)
    accuracy_report("256-tree batch mix fast (every 16th row)", g, o64)
    rate, drel, flagged = compare_fp64(g, o64)
    print(f"agree(unflagged)={rate:.6f} depth_rel_max={drel:.2e} flagged={flagged:.4%}")
    assert rate >= AGREE_MIN and drel <= DEPTH_REL_TOL
    for sc, *_ in specs:
        sc.close()


def test_fast_vs_fp64_stress_4k(gpu):
    """The stress scene (8191 nodes) at 3840x2160, one ray per pixel centre, in
    fast mode against the FP64 oracle on every 16th row (518 k rays): hit ids
    on unflagged rays and depth within 1e-4.  The reference's own FP32
    arithmetic (the FP32 oracle, which exact mode equals bit for bit) is only
    reported: its |OC|^2 - (OC.V)^2 cancels on thin cylinders, which moves
    depths by up to 3e-4 (near-tangent side exits seen almost along the axis)
    and flips about 0.07% of the sample's hits on rays that are not near a
    tangent; fast mode's closest-approach form does neither."""
    from csgray import scenes as S
    _label("stress 3840x2160", 1)
    sname, cam, light, phong, cam6, l3, ph8 = setups()[1]
    text = S.stress(4095, seed=4095)
    g = gpu_render(text, cam, light, phong, 3840, 2160, C.MODE_FAST)
    rows = (0, 2160, 16)
    oracle = O.OracleScene(text)
    o64 = oracle.render(cam6, l3, ph8, 3840, 2160, bits=64, degen=True, rows=rows)
    o32 = oracle.render(cam6, l3, ph8, 3840, 2160, bits=32, rows=rows)
    accuracy_report("stress 3840x2160 fast (every 16th row)", g, o64, rows=rows)
    accuracy_report("stress 3840x2160 reference FP32 arithmetic (every 16th row)", o32, o64, rows=rows)
    rate, drel, flagged = compare_fp64(g, o64, rows=rows)
    rate32, drel32, _ = compare_fp64(o32, o64, rows=rows)
    print(f"agree(unflagged)={rate:.6f} depth_rel_max={drel:.2e} flagged={flagged:.4%} "
          f"(the reference's FP32 arithmetic: agree(unflagged)={rate32:.6f} depth_rel_max={drel32:.2e})")
    assert rate >= AGREE_MIN
    assert drel <= DEPTH_REL_TOL


def test_fast_mode_culls_like_the_reference_on_axis_parallel_rays(gpu):
    """At an even width, the default camera's column w/2 - 1 has a direction
    with x exactly 0 (u = 1/2) from an origin with x = 0.  A tilted cylinder
    whose box (centre +- max(h/2, r), not conservative for a tilted axis)
    lies wholly at x > 0 while the cylinder itself reaches x < 0 is culled
    by the reference's slab ((l - 0) / 0 = +inf) for those rays although they
    pass through the cylinder; fast mode's reciprocal-slab boxes must cull
    them too (a true infinite reciprocal would make -o/d NaN and ignore the
    axis)."""
    from common import DEFAULT_CAM6, DEFAULT_PHONG, default_light3
    _label("axis-parallel rays, tilted cylinder outside its box", 0)
    text = "Union\nCylinder 1.2 0 0 FF8000 1 2 0 0 45\nSphere 0 30 0 00FF00 0.1\n"
    w, h = 160, 90
    col = w // 2 - 1
    sname, cam, light, phong, cam6, l3, ph8 = setups()[0]
    o = O.OracleScene(text)
    geo = o.classify(cam6, w, h)
    ref = o.render(cam6, l3, ph8, w, h, bits=64)
    # the geometry is there on that column, and the reference's box culls it
    through = (geo["status"][:, col] == 1) & (geo["prim_id"][:, col] == 0)
    assert through.sum() >= 3 and (ref["prim_id"][:, col][through] == -1).all()
    for mode in (C.MODE_FAST, C.MODE_EXACT):
        g = gpu_render(text, cam, light, phong, w, h, mode)
        assert np.array_equal(g["prim_id"][:, col], ref["prim_id"][:, col]), mode


@pytest.mark.parametrize("setup", [0, 1])
@pytest.mark.parametrize("name", scene_names() + ["cheese512", "stress1023", "tilted"])
def test_fast_tight_boxes_change_no_unflagged_ray(gpu, name, setup):
    """Fast mode culls with boxes tight around each subtree's solid where
    the reference's culling allows (a tilted cylinder that pokes out of its
    box keeps the reference's): every ray the FP64 oracle does not flag gets
    the same hit id, flags and depth bits as with the reference's boxes, and
    the per-lane kernel runs no more steps."""
    from csgray import scenes as S
    sname, cam, light, phong, cam6, l3, ph8 = setups()[setup]
    if name == "cheese512":
        text = S.cheese(512, seed=512)
    elif name == "stress1023":
        text = S.stress(1023, seed=1023)
    elif name == "tilted":  # a tilted cylinder clipped by its box beside spheres, under an Intersection
        text = ("Union\n Cylinder 0 0 0 FF0000 1 2 45 0 0\n Intersection\n  Sphere 2 0 0 00FF00 1\n"
                "  Sphere 2.5 0 0 0000FF 1\n")
    else:
        text = scene_text(name)
    w, h = (480, 270) if name in ("cheese512", "stress1023") else (160, 90)
    print(f"\n[case] {name} setup={setup} tight vs reference boxes", flush=True)
    s = C.Scene(text)
    inside = s.info().prims_in_boxes
    t = {k: v.cpu().numpy() for k, v in s.render(cam, light, phong, w, h, counters=True, coop=-1).items()}
    r = {k: v.cpu().numpy() for k, v in s.render(cam, light, phong, w, h, counters=True, coop=-1,
                                                 tight_boxes=-1).items()}
    s.close()
    o = O.OracleScene(text).render(cam6, l3, ph8, w, h, bits=64, degen=True)
    ok = o["degen"] == 0
    for k in ("prim_id", "flags"):
        assert np.array_equal(t[k][ok], r[k][ok]), k
    assert np.array_equal(t["depth"][ok].view(np.uint32), r["depth"][ok].view(np.uint32))
    steps_t = (t["counters"][..., 0] + t["counters"][..., 1]).sum()
    steps_r = (r["counters"][..., 0] + r["counters"][..., 1]).sum()
    print(f"prims_in_boxes={inside} steps tight/reference={steps_t}/{steps_r} flagged={1 - ok.mean():.4%}")
    assert steps_t <= steps_r


def test_launch_plan(gpu):
    s = C.Scene(scene_text("testCheese512.txt"))
    auto = s.plan()
    print(f"\nCUs={auto.compute_units} lds_limit={auto.lds_limit} block={auto.block} grid={auto.grid} "
          f"lds={auto.lds_bytes} scene_in_lds={auto.scene_in_lds}")
    assert auto.compute_units > 0 and auto.grid >= auto.compute_units
    g = s.plan(placement=C.PLACE_GLOBAL)
    assert g.scene_in_lds == 0
    lds = s.plan(placement=C.PLACE_LDS)
    assert lds.scene_in_lds == 1 and lds.lds_bytes <= lds.lds_limit
    # the top of the tree cached in LDS: as many of its 512 operator records as fit beside the stacks
    for block in (1024, 512, 256):
        c = s.plan(placement=C.PLACE_CACHE, block=block)
        assert c.scene_in_lds == 0 and c.block == block and c.lds_bytes <= c.lds_limit
        assert 0 < c.cache_ops <= 512 and c.lds_bytes == 10 * block * 8 + 64 * c.cache_ops
        if block == 1024:  # one workgroup a CU: all the LDS beside the stacks
            assert c.cache_ops == min(512, (c.lds_limit - 10 * block * 8 - 256) // 64)
    # the largest workgroup whose stacks and scene copy fit, and the default tail grace of a 1920x1080
    # frame (240 x 135 tiles): 950 % x waves / tiles with the state machine in the cooperative kernel
    # (exact mode), 95000 % x (waves / tiles)^5 (at least 0.5 %, at most 100 %) where it classifies by
    # intervals (fast mode, primitives inside their boxes), with budgets of 384 and 64 steps
    def lists_grace(plan):
        r = plan.grid * (plan.block // 64) / (240 * 135)
        return (min(1000, max(5, int(950000 * r ** 5 + 0.5))) + 5) // 10
    assert auto.scene_in_lds == 1 and auto.block == 1024 and auto.coop_grace == 3 and auto.coop_budget == 64
    ex = s.plan(mode=C.MODE_EXACT)
    assert ex.coop_grace == 120 and ex.coop_budget == 384
    for plan in (auto, s.plan(block=512), g):
        assert plan.coop_grace == lists_grace(plan)
    for plan in (ex, s.plan(mode=C.MODE_EXACT, block=512)):
        assert plan.coop_grace == round(950 * plan.grid * (plan.block // 64) / (240 * 135))
    assert s.plan(block=512).coop_grace == 1 and s.plan(mode=C.MODE_EXACT, block=512).coop_grace == 60
    assert s.plan(coop_grace=30).coop_grace == 30 and s.plan(coop=-1).coop_grace == 0
    # 1024-lane workgroups only for scenes of 256 operators or more: the 64-sphere cheese (64 operators)
    # and the Wikipedia scene (7) take 512 lanes, a 255-sphere cheese 512, a 256-sphere one 1024
    from csgray import scenes as S
    for text, block in ((S.cheese(64, seed=64), 512), (scene_text("testWikipedia.txt"), 512),
                        (S.cheese(255, seed=255), 512), (S.cheese(256, seed=256), 1024)):
        sc = C.Scene(text)
        assert sc.plan().block == block
        sc.close()


def test_queue_reuse_across_sizes_scenes_and_streams(gpu):
    """Back-to-back renders of different sizes and scenes on one stream, and on
    a second stream, each equal to a fresh render of the same job."""
    sname, cam, light, phong, *_ = setups()[1]
    scenes = [C.Scene(scene_text(n)) for n in ("testCheese256.txt", "testWikipedia.txt", "test5.txt")]
    sizes = [(97, 61), (256, 128), (33, 200)]
    first = [[s.render(cam, light, phong, w, h) for (w, h) in sizes] for s in scenes]
    torch.cuda.synchronize()
    side = torch.cuda.Stream()
    with torch.cuda.stream(side):
        second = [[s.render(cam, light, phong, w, h) for (w, h) in sizes] for s in scenes]
    torch.cuda.synchronize()
    again = [[s.render(cam, light, phong, w, h) for (w, h) in sizes] for s in scenes]
    torch.cuda.synchronize()
    for a_s, b_s, c_s in zip(first, second, again):
        for a, b, c in zip(a_s, b_s, c_s):
            for k in a:
                assert torch.equal(a[k], b[k]) and torch.equal(a[k], c[k]), k


@pytest.mark.parametrize("block", [256, 1024])
def test_launches_smaller_than_the_queue_stripes(gpu, block):
    """Frames of 1 to 6 tiles (fewer tiles, and with 256-lane workgroups fewer
    waves, than the work queue has stripes), each rendered twice in a row
    (the counters must return to zero after every launch), with and without
    every ray going through the cooperative kernel: bit-identical to the
    oracle."""
    print(f"\n[case] tiny frames, block={block}", flush=True)
    sname, cam, light, phong, cam6, l3, ph8 = setups()[1]
    text = scene_text("testCheese256.txt")
    s = C.Scene(text)
    o = O.OracleScene(text)
    for w, h in ((2, 2), (5, 3), (8, 8), (17, 9), (9, 17)):
        want = o.render(cam6, l3, ph8, w, h, bits=32, counters=True)
        for coop, when in ((0, 0), (1, 1), (1, 0)):
            for _ in range(2):
                g = s.render(cam, light, phong, w, h, mode=C.MODE_EXACT, counters=True, block=block,
                             coop=coop, coop_when=when)
                g = {k: v.cpu().numpy() for k, v in g.items()}
                assert np.array_equal(g["prim_id"], want["prim_id"]), (w, h, coop, when)
                assert np.array_equal(g["depth"].view(np.uint32), want["depth"].view(np.uint32)), (w, h)
                assert np.array_equal(g["flags"], want["flags"]), (w, h)
                assert np.array_equal(g["counters"], want["counters"]), (w, h)
    s.close()


def test_render_is_deterministic_and_stream_ordered(gpu):
    sname, cam, light, phong, cam6, l3, ph8 = setups()[1]
    s = C.Scene(scene_text("testCheese256.txt"))
    a = s.render(cam, light, phong, 257, 131)
    b = s.render(cam, light, phong, 257, 131)
    torch.cuda.synchronize()
    for k in a:
        assert torch.equal(a[k], b[k]), k


def test_null_outputs_allowed(gpu):
    sname, cam, light, phong, cam6, l3, ph8 = setups()[0]
    s = C.Scene(scene_text("test5.txt"))
    full = s.render(cam, light, phong, 64, 48)
    part = s.render(cam, light, phong, 64, 48, outputs={"prim_id": torch.empty((48, 64), dtype=torch.int32,
                                                                                  device="cuda")})
    torch.cuda.synchronize()
    assert torch.equal(full["prim_id"], part["prim_id"])


def test_render_argument_errors(gpu):
    sname, cam, light, phong, *_ = setups()[0]
    s = C.Scene(scene_text("test1.txt"))
    for w, h in ((1, 10), (10, 1), (0, 0)):
        with pytest.raises(C.CsgError) as e:
            s.render(cam, light, phong, w, h)
        assert e.value.status == C.E_ARG
    host = C.Scene(scene_text("test1.txt"), upload=False)
    with pytest.raises(C.CsgError):
        host.render(cam, light, phong, 16, 16)
    host.upload()
    out = host.render(cam, light, phong, 16, 16)
    torch.cuda.synchronize()
    assert out["prim_id"].shape == (16, 16)


@pytest.mark.parametrize("seed", range(30))
def test_random_tree_exact_matches_fp32_oracle(gpu, seed):
    """Random trees of the batched mix (2-1024 leaves, all three primitive
    kinds, random operators, balanced) from random orbit cameras, rendered
    alone in exact mode: bit-identical to the FP32 oracle, counters included
    (rays of heavy trees also pass through the cooperative kernel)."""
    from csgray import scenes as S
    rng = np.random.default_rng(1000 + seed)
    text = S.random_tree(rng)
    eye = S.orbit_eye(rng, 4.0, 80.0)
    _label(f"random tree seed {seed}", "orbit")
    cam = C.camera_look_at(eye, (0, 0, 0), 60)
    light, phong = C.light(S.LIGHT_TOWARD), C.phong(*S.PHONG)
    w, h = 96, 72
    g = gpu_render(text, cam, light, phong, w, h, C.MODE_EXACT, counters=True)
    o = O.OracleScene(text).render(C.cam_tuple(cam), S.LIGHT_TOWARD, O.phong8(*S.PHONG), w, h, bits=32,
                                   counters=True)
    assert np.array_equal(g["prim_id"], o["prim_id"])
    assert np.array_equal(g["flags"], o["flags"])
    assert np.array_equal(g["depth"].view(np.uint32), o["depth"].view(np.uint32))
    assert np.array_equal(g["counters"], o["counters"])


# fast vs exact RGBA on the same hit: half an 8-bit display step.  The difference is set by fast
# mode's hit-point rounding (~1e-6 relative), which moves a small sphere's normal by about
# depth error x t / r (1e-4 for r = 0.05 at t = 5) and is amplified by the specular power; the
# hardware rsq/rcp/exp2/log2 of fast shading add ~1e-6.  A wrong power at x = 0 or a missing
# normalisation shows as ~0.1.
FAST_RGBA_TOL = 2e-3


@pytest.mark.parametrize("setup", [0, 1])
@pytest.mark.parametrize("name", scene_names())
def test_fast_shading_close_to_exact(gpu, name, setup):
    """Fast mode normalises with the hardware reciprocal square root and
    takes the specular power as exp2(n log2 x); on every ray where fast and
    exact mode hit the same primitive with the same flags, RGBA agrees within
    FAST_RGBA_TOL, and misses carry the background exactly."""
    _label(name, f"{setup} fast vs exact rgba")
    sname, cam, light, phong, cam6, l3, ph8 = setups()[setup]
    text = scene_text(name)
    f = gpu_render(text, cam, light, phong, 160, 90, C.MODE_FAST)
    e = gpu_render(text, cam, light, phong, 160, 90, C.MODE_EXACT)
    same = (f["prim_id"] == e["prim_id"]) & (f["flags"] == e["flags"])
    assert same.mean() > 0.99
    d = np.abs(f["rgba"] - e["rgba"])[same]
    print(f"max |rgba fast - exact| over {int(same.sum())} rays: {d.max():.2e}")
    assert d.max() <= FAST_RGBA_TOL
    miss = same & (f["prim_id"] < 0)
    assert np.array_equal(f["rgba"][miss], e["rgba"][miss])
