"""Map ingestion on the host (no GPU): the library's PNG decoder across the
standard PNG formats, its error reporting, and the ROS map frame's yaw.

The PNG decoder stands in for the reference's lodepng_decode32_file in
OMap(filename, threshold) (RangeLib.h:158-201).  The six reference maps are
pinned against the compiled reference's own decoding in
test_oracle.py::test_png_decode_matches_reference; here each format the
decoder handles is pinned against PIL's decoding of the same file (8-bit RGBA
samples; 16-bit samples keep their high byte), including Adam7 interlacing
written by an independent encoder below.
"""
import ctypes as C
import math
import os
import struct
import zlib

import numpy as np
import pytest

import range_libc_amd as rl
from range_libc_amd import _lib


def _occupancy(path, thr):
    h = _lib.lib()
    w, hh = C.c_int(), C.c_int()
    rc = h.rl_png_occupancy(path.encode(), thr, None, C.byref(w), C.byref(hh))
    if rc != 0:
        return rc, None
    occ = np.empty((w.value, hh.value), np.uint8)
    rc = h.rl_png_occupancy(path.encode(), thr, occ.ctypes.data, C.byref(w), C.byref(hh))
    return rc, occ


def _expected_occ(rgb8, thr):
    """OMap's rule on 8-bit (R, G, B) samples: the bytes are read back to front,
    so gray = 0.229 B + 0.587 G + 0.114 R (double), as float, truncated."""
    r, g, b = (rgb8[..., 2].astype(np.float64), rgb8[..., 1].astype(np.float64), rgb8[..., 0].astype(np.float64))
    gray = (0.229 * r + 0.587 * g + 0.114 * b).astype(np.float32).astype(np.int64)
    return np.ascontiguousarray((gray.astype(np.float32) < np.float32(thr)).T.astype(np.uint8))


def _pil_cases(tmp_path):
    from PIL import Image

    rng = np.random.default_rng(7)
    H, W = 37, 53  # odd sizes: partial bytes at low bit depths
    rgb = rng.integers(0, 256, (H, W, 3), dtype=np.uint8)
    rgb[:, :10] = rgb[:, :10] // 64 * 64  # runs that make the filters' predictions matter
    cases = []

    def add(name, im, rgb8, **kw):
        p = str(tmp_path / f"{name}.png")
        im.save(p, **kw)
        cases.append((name, p, rgb8))

    add("rgb", Image.fromarray(rgb, "RGB"), rgb)
    rgba = np.concatenate([rgb, rng.integers(0, 256, (H, W, 1), dtype=np.uint8)], -1)
    add("rgba", Image.fromarray(rgba, "RGBA"), rgb)
    gray = rng.integers(0, 256, (H, W), dtype=np.uint8)
    add("gray8", Image.fromarray(gray, "L"), np.repeat(gray[..., None], 3, -1))
    la = np.stack([gray, rng.integers(0, 256, (H, W), dtype=np.uint8)], -1)
    add("gray_alpha", Image.fromarray(la, "LA"), np.repeat(gray[..., None], 3, -1))
    bits = rng.random((H, W)) < 0.4
    add("gray1", Image.fromarray(bits), np.repeat(np.where(bits, 255, 0).astype(np.uint8)[..., None], 3, -1))
    pal = rng.integers(0, 256, (16, 3), dtype=np.uint8)
    idx = rng.integers(0, 16, (H, W), dtype=np.uint8)
    pim = Image.fromarray(idx, "P")
    pim.putpalette(pal.reshape(-1).tolist())
    add("palette8", pim, pal[idx])
    add("palette4", pim, pal[idx], bits=4)
    g16 = rng.integers(0, 65536, (H, W), dtype=np.uint16)
    add("gray16", Image.fromarray(g16.astype(np.int32), "I").convert("I;16"), np.repeat((g16 >> 8).astype(np.uint8)[..., None], 3, -1))
    return cases


def test_png_formats_match_pil(tmp_path):
    for name, path, rgb8 in _pil_cases(tmp_path):
        for thr in (128.0, 77.5, 1.0, 256.0):
            rc, occ = _occupancy(path, thr)
            assert rc == 0, (name, _lib.lib().rl_last_error())
            assert np.array_equal(occ, _expected_occ(rgb8, thr)), (name, thr)


def _chunk(t, data):
    return struct.pack(">I", len(data)) + t + data + struct.pack(">I", zlib.crc32(t + data) & 0xFFFFFFFF)


def _write_adam7_rgb(path, rgb):
    """An interlaced 8-bit RGB PNG: the seven Adam7 passes, each scanline with
    the Sub filter (an encoder independent of the decoder under test)."""
    H, W, _ = rgb.shape
    raw = bytearray()
    for x0, y0, dx, dy in ((0, 0, 8, 8), (4, 0, 8, 8), (0, 4, 4, 8), (2, 0, 4, 4), (0, 2, 2, 4), (1, 0, 2, 2),
                           (0, 1, 1, 2)):
        sub = rgb[y0::dy, x0::dx]
        if sub.size == 0:
            continue
        for row in sub:
            line = row.reshape(-1).astype(np.int64)
            filt = line.copy()
            filt[3:] = (line[3:] - line[:-3]) % 256
            raw += bytes([1]) + bytes(filt.astype(np.uint8))
    ihdr = struct.pack(">IIBBBBB", W, H, 8, 2, 0, 0, 1)
    with open(path, "wb") as f:
        f.write(b"\x89PNG\r\n\x1a\n" + _chunk(b"IHDR", ihdr) + _chunk(b"IDAT", zlib.compress(bytes(raw)))
                + _chunk(b"IEND", b""))


def test_png_adam7_interlaced(tmp_path):
    from PIL import Image

    rgb = np.random.default_rng(3).integers(0, 256, (29, 19, 3), dtype=np.uint8)
    path = str(tmp_path / "adam7.png")
    _write_adam7_rgb(path, rgb)
    assert np.array_equal(np.asarray(Image.open(path).convert("RGB")), rgb)  # the encoder is valid
    for thr in (128.0, 200.0):
        rc, occ = _occupancy(path, thr)
        assert rc == 0 and np.array_equal(occ, _expected_occ(rgb, thr))


def test_png_errors_are_reported(tmp_path):
    from PIL import Image

    path = str(tmp_path / "ok.png")
    Image.fromarray(np.zeros((4, 4), np.uint8), "L").save(path)
    data = bytearray(open(path, "rb").read())
    data[20] ^= 0x01  # inside IHDR: its CRC no longer matches
    bad = tmp_path / "crc.png"
    bad.write_bytes(bytes(data))
    for p in (str(bad), str(tmp_path / "missing.png"), __file__):
        rc, _ = _occupancy(p, 128.0)
        assert rc == 4, p  # RL_ERR_IO
        assert _lib.lib().rl_last_error()


def test_png_save_then_load(tmp_path):
    img = np.random.default_rng(5).random((23, 41)) < 0.3
    m = rl.PyOMap(img.astype(np.uint8))
    p = str(tmp_path / "saved.png")
    assert m.save(p) is False
    rc, occ = _occupancy(p, 128.0)
    assert rc == 0 and np.array_equal(occ, m.occ)


# --------------------------------------------------------------------------- ROS map frame

def _q(x, y, z, w):
    from types import SimpleNamespace as NS

    return NS(x=x, y=y, z=z, w=w)


def test_ros_yaw_closed_forms():
    """rl.ros_yaw restates tf.transformations.euler_from_quaternion (axes
    'sxyz'), the reference's quaternion_to_angle (RangeLibc.pyx:120-128); tf
    is not installed, so it is pinned by what the rotation fixes: a rotation
    by theta about z has yaw theta; scaling the quaternion or flipping its
    sign changes nothing; yaw agrees with the closed form
    atan2(2(wz + xy), 1 - 2(y^2 + z^2)) for any unit quaternion; the identity
    and the degenerate quaternion give 0; at the gimbal lock yaw is 0."""
    for th in np.linspace(-math.pi + 1e-9, math.pi - 1e-9, 2001):
        q = _q(0.0, 0.0, math.sin(th / 2), math.cos(th / 2))
        assert abs(rl.ros_yaw(q) - th) < 4e-15
        for k in (1e-3, -1.0, 7.5):
            qs = _q(0.0, 0.0, k * math.sin(th / 2), k * math.cos(th / 2))
            assert abs(rl.ros_yaw(qs) - rl.ros_yaw(q)) < 4e-15
    rng = np.random.default_rng(2)
    for v in rng.normal(size=(2000, 4)):
        v /= np.linalg.norm(v)
        x, y, z, w = (float(a) for a in v)
        want = math.atan2(2 * (w * z + x * y), 1 - 2 * (y * y + z * z))
        got = rl.ros_yaw(_q(x, y, z, w))
        assert abs(math.remainder(got - want, 2 * math.pi)) < 1e-9
    assert rl.ros_yaw(_q(0.0, 0.0, 0.0, 1.0)) == 0.0
    assert rl.ros_yaw(_q(0.0, 0.0, 0.0, 0.0)) == 0.0
    s = math.sqrt(0.5)
    assert rl.ros_yaw(_q(0.0, s, 0.0, s)) == 0.0  # pitch +90 degrees
