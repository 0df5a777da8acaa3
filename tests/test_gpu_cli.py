"""The csgrender command line on the GPU.

A render of the Wikipedia scene in exact mode is checked against the FP32
oracle through the files the tool writes: the raw primitive-id, depth and
flag dumps bit for bit, the PNG within one 8-bit step (device vs host powf).
A three-frame script replays the interactive controls; each frame's dumps
must equal the oracle's render from the camera the same inputs give.
"""
import os
import subprocess
import zlib
import struct

import numpy as np
import pytest

import csgray as C
import oracle as O
from common import DEFAULT_CAM6, DEFAULT_PHONG, default_light3
from conftest import PKG, SCENE_DIR

pytestmark = pytest.mark.gpu

CLI = os.path.join(PKG, "bin", "csgrender")


def _png_pixels(path):
    data = open(path, "rb").read()
    pos, idat, w, h = 8, b"", 0, 0
    while pos < len(data):
        (n,) = struct.unpack(">I", data[pos:pos + 4])
        typ, body = data[pos + 4:pos + 8], data[pos + 8:pos + 8 + n]
        if typ == b"IHDR":
            w, h = struct.unpack(">II", body[:8])
        elif typ == b"IDAT":
            idat += body
        pos += 12 + n
    rows = np.frombuffer(zlib.decompress(idat), np.uint8).reshape(h, 1 + 4 * w)
    return rows[:, 1:].reshape(h, w, 4)


def _run(args, tmp_path):
    print("\n[case] csgrender " + " ".join(args), flush=True)
    r = subprocess.run([CLI] + args, cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    return r


def _oracle(cam6, w, h):
    text = open(os.path.join(SCENE_DIR, "testWikipedia.txt")).read()
    return O.OracleScene(text).render(cam6, default_light3(), O.phong8(*DEFAULT_PHONG), w, h, bits=32)


def test_cli_render_matches_oracle(gpu, tmp_path):
    assert os.path.exists(CLI), "build the package first (make)"
    w, h = 160, 90
    _run([os.path.join(SCENE_DIR, "testWikipedia.txt"), "--size", str(w), str(h), "--mode", "exact",
          "--out", "img.png", "--depth", "d.f32", "--id", "id.i32", "--flags", "f.u8"], tmp_path)
    o = _oracle(DEFAULT_CAM6, w, h)
    ids = np.fromfile(tmp_path / "id.i32", np.int32).reshape(h, w)
    depth = np.fromfile(tmp_path / "d.f32", np.float32).reshape(h, w)
    flags = np.fromfile(tmp_path / "f.u8", np.uint8).reshape(h, w)
    assert np.array_equal(ids, o["prim_id"])
    assert np.array_equal(depth.view(np.uint32), o["depth"].view(np.uint32))
    assert np.array_equal(flags, o["flags"])
    png = _png_pixels(tmp_path / "img.png")[::-1].astype(int)   # back to render row order
    want = np.clip(np.floor(o["rgba"] * 255 + 0.5), 0, 255).astype(int)
    assert np.abs(png - want).max() <= 1


def test_cli_script_replays_inputs(gpu, tmp_path):
    w, h = 96, 54
    script = tmp_path / "fly.txt"
    script.write_text("forward\nforward right mouse 12 -8\n\nup left mouse -30 4\n")
    _run([os.path.join(SCENE_DIR, "testWikipedia.txt"), "--size", str(w), str(h), "--mode", "exact",
          "--script", str(script), "--frames", "f_%02d.ppm", "--id", "id.i32", "--depth", "d.f32"], tmp_path)
    # the dumps hold the last frame; replay the same inputs on the camera model
    cam = C.camera(DEFAULT_CAM6[:3], DEFAULT_CAM6[3], DEFAULT_CAM6[4], DEFAULT_CAM6[5])
    C.camera_input(cam, C.KEY_FORWARD)
    C.camera_input(cam, C.KEY_FORWARD | C.KEY_RIGHT, mouse=(12.0, -8.0))
    C.camera_input(cam, 0)
    C.camera_input(cam, C.KEY_UP | C.KEY_LEFT, mouse=(-30.0, 4.0))
    for k in range(4):
        assert (tmp_path / f"f_{k:02d}.ppm").exists()
    o = _oracle(C.cam_tuple(cam), w, h)
    ids = np.fromfile(tmp_path / "id.i32", np.int32).reshape(h, w)
    depth = np.fromfile(tmp_path / "d.f32", np.float32).reshape(h, w)
    assert np.array_equal(ids, o["prim_id"])
    assert np.array_equal(depth.view(np.uint32), o["depth"].view(np.uint32))


def test_cli_reports_parse_errors(gpu, tmp_path):
    bad = tmp_path / "bad.txt"
    bad.write_text("Union\nSphere 0 0 0 FF00 1\n")
    r = subprocess.run([CLI, str(bad)], cwd=tmp_path, capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "cannot load tree" in r.stderr
