"""Pins for the shading half of the oracle: hit details and Phong
(reference: RaycastingKernels.cu:49-111 LightningKernel, :183-200 / :338-372 /
:436-457 the hit details, Float3Utils.cuh:33-44 normalize / reflect).

Two kinds of check, neither of which reads the oracle's normal or its
Flip/Exit handling:

* Closed forms at the centre pixel of a 640x360 frame from the default
  camera, whose ray is exactly (0, 0, -1): a sphere, a cube face, a cylinder
  side, both cylinder caps, a surface carved by a Difference (seen through the
  subtracted sphere's far wall), the inside of a sphere seen from within, and
  a miss.  The hit point and its normal are known exactly, and the light is
  placed where N.L and V.R take chosen values (1, 0.8, 0, -1), so the colour
  is a closed expression in ka, kd, ks, n and the primitive's colour.
* Every pixel of small frames of several scenes: the hit point comes from
  the FP64 interval classifier (no state machine), the normal from the
  primitive's geometry at that point, turned to face the viewer — the visible
  surface of a solid always faces the camera, whether the ray enters the solid
  or, from inside, leaves it, and whatever chain of Flip and Exit bits the
  combine steps set on the way.

Each of eight planted shading mistakes fails at least one of these.

A missing normalisation inside reflect() alone cannot be planted: every
normal reaching it is already of unit length (spheres, cylinder sides and cube
faces are normalised in the hit details, caps take the unit axis), so the
reference's extra normalize() only rounds.  The planted mistake 16 drops the
sphere's own normalisation instead.
"""
import math

import numpy as np
import pytest

import oracle as O
from common import DEFAULT_CAM6

W, H = 640, 360
CX, CY = 319, 179  # u = v = 0.5 exactly: the ray is the camera's forward vector
PHONG = (0.1, 0.3, 0.25, 8.0)
BG = (0.08, 0.08, 0.11, 1.0)
PH8 = O.phong8(*PHONG, bg=BG)


def colour(hexrgb):
    return np.array([int(hexrgb[i:i + 2], 16) / 255.0 for i in (0, 2, 4)])


def phong_rgba(col, N, view, L, phong=PHONG):
    """Phong with specular tinted by the object colour, clamped to [0, 1]:
    colour * (ka + kd max(N.L, 0) + ks max(V.R, 0)^n), R the mirror image of
    the incoming light direction -L about N."""
    ka, kd, ks, n = phong
    L = np.asarray(L, float) / np.linalg.norm(L)
    R = -L + 2.0 * np.dot(L, N) * N
    diff = max(float(np.dot(N, L)), 0.0)
    spec = max(float(np.dot(view, R)), 0.0) ** n
    rgb = np.clip(col * (ka + kd * diff + ks * spec), 0.0, 1.0)
    return np.array([rgb[0], rgb[1], rgb[2], 1.0])


# scene text, camera (cam6), depth, the normal facing the viewer, the visible colour
Z = np.array([0.0, 0.0, 1.0])
CASES = {
    "sphere": ("Sphere 0 0 0 FF8040 2", DEFAULT_CAM6, 3.0, Z, "FF8040"),
    "cube face": ("Cube 0 0 0 40FF80 2", DEFAULT_CAM6, 4.0, Z, "40FF80"),
    "cylinder side": ("Cylinder 0 0 0 8040FF 2 4 0 0 0", DEFAULT_CAM6, 3.0, Z, "8040FF"),
    # axis (0, 0, 1): the ray meets the top cap (Flag2, normal +V)
    "cylinder top cap": ("Cylinder 0 0 0 C0C020 1 2 90 0 0", DEFAULT_CAM6, 4.0, Z, "C0C020"),
    # axis (0, 0, -1): the cap facing the camera is the bottom one (Flag1, normal -V)
    "cylinder bottom cap": ("Cylinder 0 0 0 20C0C0 1 2 270 0 0", DEFAULT_CAM6, 4.0, Z, "20C0C0"),
    # the subtracted sphere removes the cube's front face; the ray leaves the
    # sphere at z = -0.2 inside the cube: the carved wall, Flip set by FlipR
    "difference carve": ("Difference\nCube 0 0 0 FFFFFF 2\nSphere 0 0 1 FF2040 1.2", DEFAULT_CAM6, 5.2, Z, "FF2040"),
    # the camera inside a sphere sees its far wall from within: Exit
    "inside sphere": ("Sphere 0 0 0 30FF30 10", DEFAULT_CAM6, 15.0, Z, "30FF30"),
}
# lights giving N.L = V.R = 1; N.L = V.R = 0.8; N.L = V.R = 0; N.L = V.R = -1
LIGHTS = [(0.0, 0.0, 1.0), (0.0, 0.6, 0.8), (0.0, 1.0, 0.0), (0.0, 0.0, -1.0)]


def centre(text, cam6, light, bits):
    r = O.OracleScene(text).render(cam6, light, PH8, W, H, bits=bits, rows=(CY, CY + 1, 1))
    return r["rgba"][CY, CX].astype(np.float64), float(r["depth"][CY, CX]), int(r["prim_id"][CY, CX])


def pin_closed_forms():
    for name, (text, cam6, t, N, hexrgb) in CASES.items():
        eye = np.array(cam6[:3], float)
        for bits, tol in ((64, 1e-7), (32, 2e-6)):
            for light in LIGHTS:
                rgba, depth, pid = centre(text, cam6, light, bits)
                if pid < 0 or abs(depth - t) > 1e-5 * t:
                    return False
                P = eye + t * np.array([0.0, 0.0, -1.0])
                view = (eye - P) / np.linalg.norm(eye - P)
                if np.abs(rgba - phong_rgba(colour(hexrgb), N, view, light)).max() > tol:
                    return False
    # a miss: the background colour
    rgba, depth, pid = centre("Sphere 0 10 0 FFFFFF 1", DEFAULT_CAM6, LIGHTS[0], 32)
    return pid == -1 and math.isinf(depth) and np.array_equal(rgba.astype(np.float32), np.float32(BG))


def test_closed_form_pixels():
    assert pin_closed_forms()


def test_closed_form_values_are_distinct():
    """The four lights separate every term: N.L = V.R = 1 gives
    colour * (ka + kd + ks), 0.8 gives ka + 0.8 kd + 0.8^n ks, 0 and -1 give
    colour * ka alone (both clamps)."""
    col = colour("FF8040")
    got = [phong_rgba(col, Z, Z, L)[:3] for L in LIGHTS]
    ka, kd, ks, n = PHONG
    assert np.allclose(got[0], col * (ka + kd + ks))
    assert np.allclose(got[1], col * (ka + 0.8 * kd + ks * 0.8 ** n))
    assert np.allclose(got[2], col * ka) and np.allclose(got[3], col * ka)


# ----------------------------------------------------- every pixel, geometry ---

def ray_dirs(cam6, w, h):
    b = O.camera_basis(cam6, bits=64)
    f, r, u, tan_half = b[0:3], b[3:6], b[6:9], b[9]
    xs = (np.arange(w) + 0.5) / (w - 1)
    ys = (np.arange(h) + 0.5) / (h - 1)
    nx = (2 * xs[None, :] - 1) * (w / h) * tan_half
    ny = (1 - 2 * ys[:, None]) * tan_half
    d = r[None, None, :] * nx[..., None] + u[None, None, :] * ny[..., None] + f[None, None, :]
    return d / np.linalg.norm(d, axis=-1, keepdims=True)


def geometric_normal(pf, ax64, k, P):
    """Outward unit normal of primitive k at the surface point P, or None
    where it is ambiguous (cube edges, cylinder rims)."""
    kind = int(pf[k, 1])
    c = pf[k, 2:5].astype(np.float64)
    if kind == 3:
        n = P - c
        return n / np.linalg.norm(n)
    if kind == 4:
        V = ax64[k]
        r, h = float(pf[k, 8]), float(pf[k, 9])
        base = c - (h / 2) * V
        m = float(np.dot(P - base, V))
        radial = P - base - m * V
        rho = float(np.linalg.norm(radial))
        on_cap = min(abs(m), abs(m - h)) <= 1e-6 * max(1.0, h)
        on_side = abs(rho - r) <= 1e-6 * max(1.0, r)
        if on_cap == on_side:
            return None
        if on_cap:
            return -V if abs(m) < abs(m - h) else V
        return radial / rho
    half = float(pf[k, 10]) / 2
    q = (P - c) / half
    order = np.argsort(-np.abs(q))
    if abs(q[order[1]]) > 0.999:
        return None
    n = np.zeros(3)
    n[order[0]] = math.copysign(1.0, q[order[0]])
    return n


def geometry_mismatches(text, cam6, light, w=96, h=54):
    o = O.OracleScene(text)
    k = o.render(cam6, light, PH8, w, h, bits=64, degen=True)
    c = o.classify(cam6, w, h)
    pf, ax64 = o.prims()
    d = ray_dirs(cam6, w, h)
    eye = np.array(cam6[:3], float)
    bad = checked = 0
    for y in range(h):
        for x in range(w):
            if k["degen"][y, x] or c["status"][y, x] < 0:
                continue
            if c["status"][y, x] == 0:
                if k["prim_id"][y, x] == -1:
                    checked += 1
                    bad += int(not np.array_equal(k["rgba"][y, x], np.float32(BG)))
                continue
            pid = int(c["prim_id"][y, x])
            if k["prim_id"][y, x] != pid:
                continue  # the reference's box quirks, tested elsewhere
            P = eye + c["t"][y, x] * d[y, x]
            n = geometric_normal(pf, ax64, pid, P)
            if n is None:
                continue
            nd = float(np.dot(n, d[y, x]))
            if abs(nd) < 1e-3:
                continue
            N = -n if nd > 0 else n  # the visible surface faces the viewer
            view = (eye - P) / np.linalg.norm(eye - P)
            want = phong_rgba(pf[pid, 5:8].astype(np.float64), N, view, light)
            checked += 1
            bad += int(np.abs(k["rgba"][y, x] - want).max() > 1e-5)
    return bad, checked


def geometry_cases():
    from common import setups
    view = setups()[1][4]
    light = (0.3, 0.8, 0.52)
    from conftest import scene_text
    inside = "Difference\nSphere 0 0 0 E0E0E0 10\nUnion\nSphere 0 0 -10 FF4020 4\nCylinder 3 0 -3 20FF40 1 4 90 0 0"
    return [
        (scene_text("testWikipedia.txt"), view, light),
        (scene_text("testWikipedia.txt"), DEFAULT_CAM6, (-0.5, 0.4, 0.77)),
        (scene_text("test5.txt"), view, light),
        (scene_text("testCubeCutEdges.txt"), view, light),
        (scene_text("testCheese256.txt"), DEFAULT_CAM6, light),
        (scene_text("testSphereCutByCubesAndCylinder.txt"), view, light),
        (scene_text("testCylinderSpheres2.txt"), view, light),
        (inside, DEFAULT_CAM6, light),
    ]


def pin_geometry(w=96, h=54, need=True):
    for text, cam6, light in geometry_cases():
        bad, checked = geometry_mismatches(text, cam6, light, w, h)
        if need:
            assert checked > 0.3 * w * h, "too few rays checked"
        if bad:
            return False
    return True


def test_every_pixel_matches_geometry():
    assert pin_geometry()


def test_camera_inside_scene_has_exit_and_flip_hits():
    """The inside scene exercises both negations: the big sphere's wall seen
    from within ends in an Exit record, the carved sphere and cylinder in Flip."""
    cases = geometry_cases()
    text, cam6, light = cases[-1]
    r = O.OracleScene(text).render(cam6, light, PH8, 96, 54, bits=64)
    f = r["flags"][r["prim_id"] >= 0]
    assert ((f & 4) != 0).sum() > 100 and ((f & 16) != 0).sum() > 100


# ------------------------------------------------------- planted mistakes ---

SHADING_PINS = [pin_closed_forms, lambda: pin_geometry(48, 27, need=False)]


@pytest.mark.parametrize("mutation", range(9, 17))
def test_each_planted_shading_mistake_fails_a_pin(mutation):
    try:
        O.set_mutation(mutation)
        failing = [i for i, p in enumerate(SHADING_PINS) if not p()]
    finally:
        O.set_mutation(0)
    assert failing, f"mutation {mutation} passed every shading pin"
