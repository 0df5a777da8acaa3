"""The C ABI library on the host: it loads without a GPU, exports every
function include/csg.h declares, and its host-side pieces (camera model,
light and shading defaults, scene info, argument checks) match the
reference's definitions and the oracle."""
import ctypes
import math
import os
import re
import subprocess

import numpy as np
import pytest

import csgray as C
import oracle as O
from common import DEFAULT_CAM6
from conftest import ROOT, scene_text

HEADER = os.path.join(ROOT, "include", "csg.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(csg_\w+)\s*\(", text)))


def test_header_declares_what_the_binding_binds():
    assert declared_functions() == sorted(C.symbols())


def test_library_exports_every_declared_symbol():
    C.lib()
    out = subprocess.run(["nm", "-D", "--defined-only", C.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    exported = set(re.findall(r"\sT\s(csg_\w+)", out))
    missing = [f for f in declared_functions() if f not in exported]
    assert not missing


def test_header_compiles_as_c():
    src = "#include \"csg.h\"\nint main(void){csg_camera c; csg_camera_default(&c); return (int)c.pos[2] - 5;}\n"
    p = subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-fsyntax-only", "-I", os.path.dirname(HEADER),
                        "-x", "c", "-"], input=src, text=True, capture_output=True)
    assert p.returncode == 0, p.stderr


def test_version_and_last_error():
    assert b"gfx950" in C.lib().csg_version()
    with pytest.raises(C.CsgError):
        C.Scene("Bogus 1 2 3", upload=False)
    assert b"Unrecognized keyword" in C.lib().csg_last_error()


def test_default_camera_matches_reference():
    cam = C.camera_default()
    assert tuple(cam.pos) == (0.0, 0.0, 5.0) and cam.pitch == 0.0 and cam.yaw == 0.0
    assert cam.vfov == np.float32(np.float32(90.0) * np.float32(3.14159) / np.float32(180.0))
    b = C.camera_basis(cam)
    assert np.allclose(b[0:3], [0, 0, -1]) and np.allclose(b[3:6], [1, 0, 0]) and np.allclose(b[6:9], [0, -1, 0])


@pytest.mark.parametrize("cam6", [DEFAULT_CAM6, (3.0, 2.5, 4.0, -0.4636476, 0.6435011, 1.0471975),
                                  (-2.0, 7.0, 1.0, 1.2, -2.0, 0.5), (0.0, 0.0, 9.0, 2.0, 0.3, 1.7)])
def test_camera_basis_bitwise_equal_to_oracle(cam6):
    mine = np.array(C.camera_basis(C.camera(cam6[:3], cam6[3], cam6[4], cam6[5])), np.float32)
    ref = O.camera_basis(cam6, bits=32).astype(np.float32)
    assert np.array_equal(mine.view(np.uint32), ref.view(np.uint32))


def test_pitch_is_clamped_like_set_rotation():
    lim = np.float32(np.float32(89.0) * np.float32(3.14159) / np.float32(180.0))
    up = C.camera_basis(C.camera((0, 0, 0), 2.0, 0.0, 1.0))
    assert abs(up[1] - math.sin(lim)) < 1e-6


def test_look_at_recovers_the_view_configuration():
    cam = C.camera_look_at((3, 2.5, 4), (0, 0, 0), 60)
    assert abs(cam.pitch - (-0.4636476)) < 1e-6 and abs(cam.yaw - 0.6435011) < 1e-6
    assert cam.vfov == np.float32(np.float32(60.0) * np.float32(3.14159) / np.float32(180.0))
    f = np.array(C.camera_basis(cam)[0:3])
    assert np.allclose(f, -np.array([3, 2.5, 4]) / math.sqrt(31.25), atol=1e-6)
    with pytest.raises(C.CsgError):
        C.camera_look_at((1, 1, 1), (1, 1, 1), 60)


def test_light_and_phong_defaults():
    l = C.light_default()
    p, a = np.float32(-60.0 * 3.14159 / 180.0), np.float32(-45.0 * 3.14159 / 180.0)
    assert np.allclose(list(l.dir), [math.sin(p) * math.cos(a), math.cos(p), math.sin(p) * math.sin(a)], atol=1e-6)
    ph = C.phong_default()
    assert (ph.ka, ph.kd, ph.ks, ph.shininess) == (np.float32(0.2), np.float32(0.8), np.float32(0.7), 30.0)
    assert np.allclose(list(ph.bg), [0.08, 0.08, 0.11, 1.0])


def test_scene_info_counts():
    s = C.Scene(scene_text("testCubeCutEdges.txt"), upload=False)
    i = s.info()
    assert (i.nodes, i.primitives, i.op_depth, i.spheres, i.cylinders, i.cubes) == (25, 13, 5, 0, 12, 1)
    assert i.device_bytes == 0


@pytest.mark.parametrize("text,inside", [
    ("Sphere 0 0 0 FF0000 1", 1),
    ("Cube 0 0 0 FF0000 1", 1),
    ("Cylinder 0 0 0 FF0000 1 2 0 0 0", 1),      # axis y: half-extents (r, h/2, r) = the box's 1
    ("Cylinder 0 0 0 FF0000 1 2 90 0 0", 1),     # axis z
    ("Cylinder 0 0 0 FF0000 0.1 4 30 0 0", 1),   # long and thin: inside the h/2 cube however tilted
    ("Cylinder 0 0 0 FF0000 1 2 45 0 0", 0),     # y and z half-extents sqrt(2) > 1
    ("Union\n Sphere 0 0 0 FF0000 1\n Cylinder 3 0 0 FF0000 1 1 0 0 30\n", 0),
])
def test_scene_info_reports_primitives_inside_their_boxes(text, inside):
    s = C.Scene(text, upload=False)
    assert s.info().prims_in_boxes == inside


S1 = "Sphere 0 0 0 FF0000 1\n"


@pytest.mark.parametrize("text,groups", [
    ("Union\n" + S1 * 2, 0),                                    # two members: stays binary
    ("Union\nUnion\n" + S1 * 3, 1),                            # three members under two Unions
    ("Intersection\nIntersection\n" + S1 * 3, 0),              # only Unions are flattened
    ("Union\nDifference\nUnion\n" + S1 + "Union\n" + S1 * 3 + S1, 1),  # the outer Union has 2 members
    ("Union\nUnion\nDifference\n" + S1 * 2 + S1 + "Union\n" + S1 * 2, 1),  # a Difference member
])
def test_scene_info_counts_union_groups(text, groups):
    assert C.Scene(text, upload=False).info().union_groups == groups


def test_union_groups_split_above_1024_members():
    from csgray import scenes as S
    assert C.Scene(S.cheese(512, seed=512), upload=False).info().union_groups == 1
    assert C.Scene(S.cheese(1024, seed=3), upload=False).info().union_groups == 1
    # 2048 members: the top Union stays binary, its two 1024-member halves are groups
    assert C.Scene(S.cheese(2048, seed=3), upload=False).info().union_groups == 2


def test_render_refuses_host_only_scene_without_touching_the_gpu():
    s = C.Scene(scene_text("test1.txt"), upload=False)
    o = C.Outputs(None, None, None, None)
    st = C.lib().csg_render(s.handle, C.camera_default(), C.light_default(), C.phong_default(), 16, 16, o, None)
    assert st == C.E_ARG


def test_null_arguments_are_errors_not_crashes():
    L = C.lib()
    assert L.csg_scene_info_get(None, None) == C.E_ARG
    assert L.csg_scene_export(None, None, None, None) == C.E_ARG
    assert L.csg_scene_free(None) == C.OK
    assert L.csg_render(None, None, None, None, 4, 4, None, None) == C.E_ARG


def test_jitter_hash_matches_oracle_and_is_uniform():
    rng = np.random.default_rng(0)
    vals = []
    for _ in range(2000):
        seed, pixel, sample, axis = (int(v) for v in rng.integers(0, 2**31, size=4))
        axis &= 1
        a = C.jitter(seed, pixel, sample, axis)
        assert a == O.jitter(seed, pixel, sample, axis)
        assert 0.0 <= a < 1.0
        vals.append(a)
    hist = np.histogram(vals, bins=10, range=(0, 1))[0]
    assert hist.min() > 140 and hist.max() < 260


@pytest.mark.parametrize("field,value", [("coop_when", 2), ("coop_when", -1), ("coop_grace", 100001),
                                         ("samples", -1), ("samples", 1025), ("mode", 7)])
def test_bad_render_options_are_rejected_before_any_device_call(field, value):
    s = C.Scene(scene_text("test1.txt"), upload=False)
    opts = C.RenderOpts(C.MODE_FAST, 0, C.PLACE_AUTO, 0, None, None, 1, 0, 0, 0, 0, 0)
    setattr(opts, field, value)
    info = C.LaunchInfo()
    st = C.lib().csg_render_plan(s.handle, ctypes.byref(opts), ctypes.byref(info))
    assert st == C.E_ARG, (field, value, st)
    s.close()
