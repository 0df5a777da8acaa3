"""Pins for the oracle, independent of the restatement's own formulas.

* Known answers measured on the unmodified reference sources (a host build of
  RaycastingKernels.cu / CSGTree.cu / Camera.cpp behind a small header shim,
  FP32, made once outside this repository, whose figures are written into
  the tests below): hit counts and mean per-ray event counts at 640x360.
  They corroborate; the checks below carry the weight.
* Closed forms: cylinder axis = Ry(y) Rx(x) Rz(z) (0,1,0); leaf boxes; camera
  basis; the centre ray's analytic depth on single primitives.
* An interval classifier that evaluates the CSG expression between the
  sorted boundary points of all primitives, with no Kensler state machine:
  on scenes whose boxes are conservative the FP64 restatement must agree with
  it on every ray it does not flag as degenerate.
* Invariants: a far-away operand leaves the image unchanged.
* Planted mistakes: each of eight plausible errors in the restatement fails
  at least one of the pins above (the shading half has its own pins and
  planted mistakes in test_oracle_shading.py).
* The degenerate-decision flagger: every ray whose FP64 result a move of the
  view by 1e-6 of its distance flips is flagged, and a tangency margin 1000
  times too small is caught.
"""
import math

import numpy as np
import pytest

import oracle as O
from common import DEFAULT_CAM6, DEFAULT_PHONG, default_light3, setups
from conftest import scene_text

LIGHT = default_light3()
PH8 = O.phong8(*DEFAULT_PHONG)


def render(name_or_text, cam6=DEFAULT_CAM6, w=640, h=360, bits=32, **kw):
    text = scene_text(name_or_text) if name_or_text.endswith(".txt") else name_or_text
    return O.OracleScene(text).render(cam6, LIGHT, PH8, w, h, bits=bits, **kw)


def view_cam6():
    return setups()[1][4]


# ----------------------------------------------------------- known answers ---

def pin_single_primitive_hits():
    got = {n: int((render(n)["prim_id"] >= 0).sum())
           for n in ("testSingleSphere.txt", "testSingleCube.txt", "testSingleCylinder.txt")}
    return got == {"testSingleSphere.txt": 19305, "testSingleCube.txt": 7921, "testSingleCylinder.txt": 44091}


def pin_wikipedia_view():
    r = render("testWikipedia.txt", cam6=view_cam6(), counters=True)
    ids = r["prim_id"]
    split = np.bincount(ids[ids >= 0], minlength=5).tolist()
    ev = r["counters"].reshape(-1, 6).mean(0)
    return (split == [5550, 4234, 1707, 2068, 2014]
            and np.allclose(ev, [2.08, 2.17, 3.94, 0.77, 0.43, 0.15], atol=0.006))


def pin_cheese512_events():
    r = render("testCheese512.txt", counters=True)
    ev = r["counters"].reshape(-1, 6).mean(0)
    hit = float((r["prim_id"] >= 0).mean())
    return np.allclose(ev, [110.3, 110.5, 220.4, 7.18, 0.0, 0.34], atol=0.06) and abs(hit - 0.235) < 0.0006


def test_known_answer_single_primitives():
    assert pin_single_primitive_hits()


def test_known_answer_wikipedia_view():
    assert pin_wikipedia_view()


def test_known_answer_cheese512():
    assert pin_cheese512_events()


# ------------------------------------------------------------ closed forms ---

def pin_centre_depths():
    # 640x360: pixel (319, 179) has u = v = 0.5 exactly, i.e. the optical axis.
    want = {"testSingleSphere.txt": 3.0, "testSingleCube.txt": 4.0, "testSingleCylinder.txt": 3.0}
    for n, t in want.items():
        for bits in (32, 64):
            if abs(float(render(n, bits=bits)["depth"][179, 319]) - t) > 1e-6:
                return False
    return True


def test_centre_ray_depths():
    assert pin_centre_depths()


def rot_axis(dx, dy, dz):
    x, y, z = (math.radians(v) for v in (dx, dy, dz))
    Rx = np.array([[1, 0, 0], [0, math.cos(x), -math.sin(x)], [0, math.sin(x), math.cos(x)]])
    Ry = np.array([[math.cos(y), 0, math.sin(y)], [0, 1, 0], [-math.sin(y), 0, math.cos(y)]])
    Rz = np.array([[math.cos(z), -math.sin(z), 0], [math.sin(z), math.cos(z), 0], [0, 0, 1]])
    return Ry @ Rx @ Rz @ np.array([0.0, 1.0, 0.0])


@pytest.mark.parametrize("rot,axis", [((0, 0, 0), (0, 1, 0)), ((90, 0, 0), (0, 0, 1)), ((0, 0, 90), (-1, 0, 0)),
                                      ((0, 90, 0), (0, 1, 0)), ((180, 0, 0), (0, -1, 0))])
def test_cylinder_axis_special_angles(rot, axis):
    o = O.OracleScene("Cylinder 0 0 0 FF0000 1 2 %g %g %g" % rot)
    _, ax64 = o.prims()
    assert np.allclose(ax64[0], axis, atol=1e-15)


def test_cylinder_axis_is_rotation_product():
    rng = np.random.default_rng(5)
    for _ in range(50):
        rot = rng.uniform(0, 360, size=3)
        o = O.OracleScene("Cylinder 0 0 0 FF0000 1 2 %.17g %.17g %.17g" % tuple(rot))
        pf, ax64 = o.prims()
        assert np.allclose(ax64[0], rot_axis(*rot), atol=1e-12)
        assert np.allclose(pf[0, 11:14], ax64[0].astype(np.float32), atol=0)


def test_leaf_and_operator_boxes():
    o = O.OracleScene("Union\n Sphere 1 2 3 FF0000 0.5\n Union\n  Cylinder -1 0 0 00FF00 0.3 4 30 40 50\n"
                      "  Cube 0 0 5 0000FF 3\n")
    links, b32, b64 = o.nodes()
    assert np.allclose(b64[1], [0, 1, 2, 2, 3, 4])                 # sphere: centre +- 2r
    assert np.allclose(b64[3], [-3, -2, -2, 1, 2, 2])              # cylinder: centre +- max(h/2, r)
    assert np.allclose(b64[4], [-1.5, -1.5, 3.5, 1.5, 1.5, 6.5])   # cube: centre +- edge/2
    assert np.allclose(b64[2], [-3, -2, -2, 1.5, 2, 6.5])          # operators: union of children
    assert np.allclose(b64[0], [-3, -2, -2, 2, 3, 6.5])


def test_camera_basis_closed_forms():
    b = O.camera_basis(DEFAULT_CAM6, bits=64)
    assert np.allclose(b[0:3], [0, 0, -1]) and np.allclose(b[3:6], [1, 0, 0]) and np.allclose(b[6:9], [0, -1, 0])
    cam6 = view_cam6()
    b = O.camera_basis(cam6, bits=64)
    f = -np.array([3.0, 2.5, 4.0]) / math.sqrt(31.25)
    assert np.allclose(b[0:3], f, atol=1e-6)
    assert abs(cam6[3] - (-0.4636476)) < 1e-6 and abs(cam6[4] - 0.6435011) < 1e-6
    assert np.allclose(np.cross(b[0:3], b[3:6]), b[6:9], atol=1e-6)
    assert abs(b[9] - math.tan(cam6[5] / 2)) < 1e-12


# -------------------------------------------------------------- classifier ---

CONSERVATIVE = ["test1.txt", "test2.txt", "test3.txt", "test4.txt", "test5.txt", "testPipe.txt",
                "testCubePipe.txt", "testWikipedia.txt", "testWikipediaLeft.txt", "testWikipediaRight.txt",
                "testWikipediaMult.txt", "testCubeCutEdges.txt", "testSphereCutByCubesAndCylinder.txt",
                "testSingleSphere.txt", "testSingleCube.txt", "testSingleCylinder.txt",
                "testCylinderSpheres2.txt", "testCheese256.txt"]


def classifier_agreement(name, cam6, w=96, h=54):
    text = scene_text(name)
    o = O.OracleScene(text)
    k = o.render(cam6, LIGHT, PH8, w, h, bits=64, degen=True)
    c = o.classify(cam6, w, h)
    use = (c["status"] >= 0) & (k["degen"] == 0)
    khit = k["prim_id"] >= 0
    chit = c["status"] == 1
    same = (khit == chit)
    both = khit & chit & use
    same &= np.where(khit & chit, k["prim_id"] == c["prim_id"], True)
    enter = (k["flags"] & 2) != 0
    same &= np.where(khit & chit, enter == (c["enter"] == 1), True)
    t_ok = np.abs(k["depth"][both].astype(np.float64) - c["t"][both]) <= 1e-5 * c["t"][both]
    return int((~same & use).sum()), int(use.sum()), bool(t_ok.all())


@pytest.mark.parametrize("setup", [0, 1])
@pytest.mark.parametrize("name", CONSERVATIVE)
def test_kensler_matches_interval_classifier(name, setup):
    bad, used, t_ok = classifier_agreement(name, setups()[setup][4])
    assert used > 0.9 * 96 * 54 * 0.5
    assert bad == 0 and t_ok


def test_tilted_cylinder_box_clips_like_the_reference():
    """The reference's cylinder box (centre +- max(h/2, r)) is not conservative
    for a tilted axis: rays that meet the cylinder only outside that box miss
    it (BVHNode.cuh:34-49).  The restatement must keep that quirk."""
    # r = 1, h = 2 tilted 45 degrees about x: the true extent in y and z is
    # sqrt(2) but the box is +-1.
    text = "Union\n Cylinder 0 0 0 FF0000 1 2 45 0 0\n Sphere 30 0 0 00FF00 0.1\n"
    o = O.OracleScene(text)
    cam6 = (6.0, 0.0, 0.0, 0.0, math.pi / 2, math.radians(60))
    w = h = 96
    k = o.render(cam6, LIGHT, PH8, w, h, bits=64)
    c = o.classify(cam6, w, h)
    clipped = (c["status"] == 1) & (k["prim_id"] < 0)
    assert clipped.sum() > 20
    # every clipped ray meets the cylinder at a point outside the box
    b = O.camera_basis(cam6, bits=64)
    ys, xs = np.nonzero(clipped)
    for y, x in zip(ys, xs):
        u, v = (x + 0.5) / (w - 1), (y + 0.5) / (h - 1)
        nx, ny = (2 * u - 1) * (w / h) * b[9], (1 - 2 * v) * b[9]
        d = b[3:6] * nx + b[6:9] * ny + b[0:3]
        d /= np.linalg.norm(d)
        p = np.array(cam6[:3]) + c["t"][y, x] * d
        assert np.abs(p).max() > 1.0
    # and where the ray does cross the box, the two agree
    inside = (c["status"] == 1) & (k["prim_id"] >= 0)
    assert (k["prim_id"][inside] == c["prim_id"][inside]).all()


# -------------------------------------------------------------- invariants ---

FAR = "Sphere 100 100 -100 00FF00 1"


@pytest.mark.parametrize("op", ["Union", "Difference"])
@pytest.mark.parametrize("name", ["testWikipedia.txt", "test5.txt", "testCubeCutEdges.txt"])
def test_far_operand_changes_nothing(name, op):
    base = scene_text(name)
    wrapped = f"{op}\n{base}\n{FAR}\n"
    a = render(base, w=160, h=90)
    b = render(wrapped, w=160, h=90)
    assert np.array_equal(a["prim_id"], b["prim_id"])
    assert np.array_equal(a["depth"].view(np.uint32), b["depth"].view(np.uint32))


def test_fp32_fp64_mismatches_are_all_flagged_on_wikipedia():
    cam6 = view_cam6()
    r32 = render("testWikipedia.txt", cam6=cam6)
    r64 = render("testWikipedia.txt", cam6=cam6, bits=64, degen=True)
    mism = r32["prim_id"] != r64["prim_id"]
    assert 0 < mism.sum() < 0.003 * mism.size     # the method's own FP32 rounding noise at coplanar faces
    assert (mism & (r64["degen"] == 0)).sum() == 0
    assert (r64["degen"] != 0).mean() < 0.05


def test_stack_capacity_matches_reference_for_shallow_trees():
    """Depth <= 15 never overflows the reference's 32-entry stacks."""
    r = render("testCheese512.txt", w=128, h=72, bits=64, degen=True)
    assert ((r["degen"] & O.DG_OVERFLOW) == 0).all()


# ------------------------------------------------------- planted mistakes ---

def pin_classifier_small():
    for name in ("testWikipedia.txt", "test5.txt", "testCubePipe.txt", "testCheese256.txt"):
        for setup in (0, 1):
            bad, _, t_ok = classifier_agreement(name, setups()[setup][4], w=64, h=36)
            if bad or not t_ok:
                return False
    return True


PINS = [pin_single_primitive_hits, pin_wikipedia_view, pin_cheese512_events, pin_centre_depths,
        pin_classifier_small]


@pytest.mark.parametrize("mutation", range(1, 9))
def test_each_planted_mistake_fails_a_pin(mutation):
    try:
        O.set_mutation(mutation)
        failing = [p.__name__ for p in PINS if not p()]
    finally:
        O.set_mutation(0)
    assert failing, f"mutation {mutation} passed every pin"


def test_pins_pass_without_mistakes():
    O.set_mutation(0)
    assert all(p() for p in PINS)


# ------------------------------------------------------------- the flagger ---

def _view_moves(cam6, rel=1e-6):
    """The camera moved by `rel` of its distance along each axis, and turned
    by `rel` radians in pitch and in yaw (each way)."""
    d = math.sqrt(sum(c * c for c in cam6[:3]))
    out = []
    for k in range(5):
        for sgn in (1, -1):
            v = list(cam6)
            v[k] = float(np.float32(v[k] + sgn * rel * (d if k < 3 else 1.0)))
            out.append(tuple(v))
    return out


def unflagged_flips(text, w, h, rel=1e-6):
    """Rays whose FP64 hit id or enter/exit changes when the view moves by
    `rel` (half the flagger's tangency margin), and how many of them the
    unmoved render leaves unflagged."""
    cam6 = view_cam6()
    o = O.OracleScene(text)
    base = o.render(cam6, LIGHT, PH8, w, h, bits=64, degen=True)
    flipped = np.zeros((h, w), bool)
    for v in _view_moves(cam6, rel):
        r = o.render(v, LIGHT, PH8, w, h, bits=64)
        flipped |= (r["prim_id"] != base["prim_id"]) | ((r["flags"] & 14) != (base["flags"] & 14))
    return int(flipped.sum()), int((flipped & (base["degen"] == 0)).sum())


def test_flags_cover_every_ray_a_small_move_of_the_view_flips():
    """The degenerate set is meant to hold the rays whose result is not fixed
    at FP32 input precision.  Moving the camera by 1e-6 of its distance (or
    turning it by 1e-6 rad) flips the coplanar ties of the Wikipedia view and
    the silhouette rays of the stress scene's small primitives; every flipped
    ray must be flagged in the unmoved render."""
    from csgray import scenes as S
    for text, w, h, least in ((scene_text("testWikipedia.txt"), 320, 180, 100),
                              (S.stress(4095, seed=4095), 640, 360, 1)):
        flips, missed = unflagged_flips(text, w, h)
        assert flips >= least and missed == 0, (flips, missed)


def test_an_under_flagging_tangency_test_is_caught():
    """Planted: a tangency margin 1000 times too small leaves silhouette flips
    of the stress scene unflagged."""
    from csgray import scenes as S
    try:
        O.set_tangent_eta(2e-9)
        flips, missed = unflagged_flips(S.stress(4095, seed=4095), 640, 360)
    finally:
        O.set_tangent_eta(2e-6)
    assert missed > 0
