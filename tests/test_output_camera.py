"""Headless output writers and the interactive camera control (CPU only).

PNG files are decoded here independently (chunk CRCs with zlib.crc32, the
IDAT stream with zlib.decompress, filter bytes checked), so a wrong row
order, channel mapping or checksum fails.  Camera control is pinned against
the closed forms of Camera.h:37-47 and Application.cpp:85-119.
"""
import math
import struct
import zlib

import numpy as np
import pytest

import csgray as C


def _expected_bytes(rgba):
    v = np.nan_to_num(rgba.astype(np.float32), nan=0.0)
    out = np.where(v <= 0, 0, np.where(v >= 1, 255, np.floor(v * np.float32(255) + np.float32(0.5))))
    return out.astype(np.uint8)[::-1]  # top row first


def _read_png(path):
    data = open(path, "rb").read()
    assert data[:8] == b"\x89PNG\r\n\x1a\n"
    pos, chunks = 8, []
    while pos < len(data):
        (n,) = struct.unpack(">I", data[pos:pos + 4])
        typ = data[pos + 4:pos + 8]
        body = data[pos + 8:pos + 8 + n]
        (crc,) = struct.unpack(">I", data[pos + 8 + n:pos + 12 + n])
        assert crc == zlib.crc32(typ + body) & 0xFFFFFFFF, typ
        chunks.append((typ, body))
        pos += 12 + n
    assert [c[0] for c in chunks] == [b"IHDR", b"IDAT", b"IEND"]
    w, h, depth, ctype, comp, filt, inter = struct.unpack(">IIBBBBB", chunks[0][1])
    assert (depth, ctype, comp, filt, inter) == (8, 6, 0, 0, 0)
    raw = zlib.decompress(chunks[1][1])
    rows = np.frombuffer(raw, np.uint8).reshape(h, 1 + 4 * w)
    assert (rows[:, 0] == 0).all()
    return rows[:, 1:].reshape(h, w, 4)


def _test_image(h, w, seed=1):
    rng = np.random.default_rng(seed)
    img = rng.uniform(-0.2, 1.2, size=(h, w, 4)).astype(np.float32)
    img[0, 0, 0] = np.nan
    img[0, 1, :] = (0.0, 1.0, 0.5, 0.25)
    return img


def test_png_roundtrip_small(tmp_path):
    img = _test_image(3, 5)
    p = tmp_path / "a.png"
    C.write_image(p, img)
    assert np.array_equal(_read_png(p), _expected_bytes(img))


def test_png_spans_several_stored_blocks(tmp_path):
    img = _test_image(200, 300, seed=2)   # 240,200 bytes of scanlines -> 4 blocks
    p = tmp_path / "b.png"
    C.write_image(p, img)
    assert np.array_equal(_read_png(p), _expected_bytes(img))


def test_ppm_header_rows_and_channels(tmp_path):
    img = _test_image(4, 6, seed=3)
    p = tmp_path / "c.ppm"
    C.write_image(p, img)
    data = open(p, "rb").read()
    head = b"P6\n6 4\n255\n"
    assert data.startswith(head)
    got = np.frombuffer(data[len(head):], np.uint8).reshape(4, 6, 3)
    assert np.array_equal(got, _expected_bytes(img)[..., :3])


@pytest.mark.parametrize("dtype", [np.float32, np.int32, np.uint8])
@pytest.mark.parametrize("flip", [False, True])
def test_raw_dump(tmp_path, dtype, flip):
    a = (np.arange(7 * 5) * 3 - 11).astype(dtype).reshape(7, 5)
    p = tmp_path / "d.raw"
    C.write_raw(p, a, flip_rows=flip)
    back = np.fromfile(p, dtype).reshape(7, 5)
    assert np.array_equal(back, a[::-1] if flip else a)


def test_writer_errors(tmp_path):
    img = _test_image(2, 2)
    with pytest.raises(C.CsgError):
        C.write_image(tmp_path / "missing_dir" / "x.png", img)
    with pytest.raises(C.CsgError):
        C.write_image(tmp_path / "x.png", np.zeros((0, 3, 4), np.float32))


def _basis(pitch, yaw):
    f = np.array([-math.sin(yaw) * math.cos(pitch), math.sin(pitch), -math.cos(yaw) * math.cos(pitch)])
    r = np.array([math.cos(yaw), 0.0, -math.sin(yaw)])
    return f / np.linalg.norm(f), r / np.linalg.norm(r)


def test_move_default_camera_axes():
    cam = C.camera((0, 0, 5), 0.0, 0.0, math.radians(90))
    C.camera_move(cam, 1.0, 0.0, 0.0)          # forward is -z
    assert tuple(cam.pos) == pytest.approx((0, 0, 4))
    C.camera_move(cam, 0.0, 2.0, 0.0)          # right is +x
    assert tuple(cam.pos) == pytest.approx((2, 0, 4))
    C.camera_move(cam, 0.0, 0.0, 0.5)          # up is world +y, not the camera's up (-y here)
    assert tuple(cam.pos) == pytest.approx((2, 0.5, 4))


def test_move_tilted_camera_closed_form():
    cam = C.camera((1, 2, 3), 0.3, 0.7, 1.0)
    f, r = _basis(0.3, 0.7)
    C.camera_move(cam, 0.4, -0.25, 0.1)
    want = np.array([1, 2, 3]) + 0.4 * f - 0.25 * r + np.array([0, 0.1, 0])
    assert np.allclose(np.array(cam.pos), want, atol=2e-6)


def test_rotate_clamps_pitch_only():
    lim = np.float32(89.0) * np.float32(3.14159) / np.float32(180.0)
    cam = C.camera((0, 0, 5), 0.0, 0.0, 1.0)
    C.camera_rotate(cam, 10.0, 7.0)
    assert cam.pitch == pytest.approx(float(lim), abs=0) and cam.yaw == pytest.approx(7.0)
    C.camera_rotate(cam, -20.0, -1.0)
    assert cam.pitch == pytest.approx(-float(lim), abs=0) and cam.yaw == pytest.approx(6.0)


def test_input_frame_rotates_then_moves():
    cam = C.camera((0, 0, 5), 0.0, 0.0, 1.0)
    C.camera_input(cam, C.KEY_FORWARD | C.KEY_BACKWARD | C.KEY_RIGHT | C.KEY_UP, mouse=(10.0, -20.0))
    yaw, pitch = -0.005 * 10.0, -0.005 * -20.0
    assert cam.yaw == pytest.approx(yaw) and cam.pitch == pytest.approx(pitch)
    f, r = _basis(pitch, yaw)
    want = np.array([0, 0, 5]) + 0.0 * f + 0.1 * r + np.array([0, 0.1, 0])
    assert np.allclose(np.array(cam.pos), want, atol=2e-6)
    # no mouse: only the move
    C.camera_input(cam, C.KEY_LEFT | C.KEY_DOWN)
    want = want - 0.1 * r - np.array([0, 0.1, 0])
    assert np.allclose(np.array(cam.pos), want, atol=4e-6)
