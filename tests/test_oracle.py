"""CPU tests that pin the oracle (oracle/rl_oracle.c) before it is trusted.

Each oracle function is checked against
  * golden vectors written by the compiled reference (tests/golden/*.npz),
  * and an independent closed form or brute force where the mathematics fixes
    the answer, chosen so that a plausible mistake (a transposed index, a
    dropped term, a wrong sign or rounding) fails.
When the compiled reference (oracle/_ref/libref.so) is present, the oracle is
also compared with it live on random inputs.
"""
import hashlib
import math
import os
import subprocess

import numpy as np
import pytest

from oracle import oracle as O
from range_libc_amd import workloads as W
from tests.conftest import GOLDEN, assert_bits_equal

FLT_MAX = np.finfo(np.float32).max


def npz(name):
    return np.load(os.path.join(GOLDEN, name), allow_pickle=False)


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def unpack(z, tag):
    shape = tuple(z[f"{tag}_shape"])
    return np.unpackbits(z[f"{tag}_occ"])[: shape[0] * shape[1]].reshape(shape)


PNG_KEYS = ["small_map", "quad_map", "single_pixel_map", "synthetic_map", "test", "basement_hallways_5cm"]


# --------------------------------------------------------------------------- generators and inputs

def test_generators_reproduce_fixture_maps():
    assert sha(W.rect_disk_map(512, seed=0)) == str(npz("rect_disk_512.npz")["map_sha"])
    assert sha(W.corridor_map(2048, seed=1)) == str(npz("corridor_2048.npz")["map_sha"])
    assert sha(W.sensor_table(501)) == str(npz("rect_disk_512.npz")["table_sha"])


@pytest.mark.parametrize("key", PNG_KEYS)
def test_png_decode_matches_reference(key):
    """PyOMap(path, threshold) reproduces OMap(filename, threshold) (RangeLib.h:159-201)."""
    import range_libc_amd as rl

    z = npz("png_maps.npz")
    path = os.path.join(GOLDEN, "maps", key.replace("_map", ".map") + ".png") if key.endswith("_map") \
        else os.path.join(GOLDEN, "maps", key + ".png")
    for thr in (128, 1):
        want = unpack(z, f"{key}_t{thr}")
        got = rl.PyOMap(path, float(thr)).occ.T
        assert np.array_equal(got, want), (key, thr)


def test_save_round_trip(tmp_path):
    """PyOMap.save writes OMap::save's image (RangeLib.h:264-291: opaque RGBA,
    black where occupied, white elsewhere), which PyOMap(path) reads back."""
    from PIL import Image

    import range_libc_amd as rl

    img = W.rect_disk_map(64, seed=3)[:, :50]  # non-square: 50 wide, 64 high
    m = rl.PyOMap(img)
    path = str(tmp_path / "map.png")
    assert m.save(path) is False
    rgba = np.asarray(Image.open(path))
    assert rgba.shape == (64, 50, 4) and np.all(rgba[..., 3] == 255)
    assert np.array_equal(rgba[..., 0] == 0, img != 0) and np.array_equal(rgba[..., :3].min(-1), rgba[..., :3].max(-1))
    assert np.array_equal(rl.PyOMap(path).occ, m.occ)


def test_png_load_failure_flags_error(tmp_path, capsys):
    """OMap(filename) on an unreadable file prints the decoder error and
    reports error() (RangeLib.h:159-168); the map left behind is 1x1 and free
    here (the reference's grid is empty)."""
    import range_libc_amd as rl

    bad = tmp_path / "not_a.png"
    bad.write_bytes(b"not a png")
    for path in (str(bad), str(tmp_path / "missing.png")):
        m = rl.PyOMap(path)
        assert m.error() and (m.width(), m.height()) == (1, 1)
    assert "ERROR" in capsys.readouterr().out
    assert not rl.PyOMap(np.zeros((3, 4), np.uint8)).error()


def test_ros_occupancy_grid_ingestion():
    """PyOMap(OccupancyGrid) as RangeLibc.pyx:146-167 builds it: OMap(height,
    width) with grid[x][y] = data[x * width + y] > 10 (x over the message's
    rows; -1 unknown and 0..10 free), and the world frame from the map origin,
    angle = -yaw of its quaternion."""
    from types import SimpleNamespace as NS

    import range_libc_amd as rl

    h, w, yaw = 5, 7, 0.4
    data = np.random.default_rng(4).choice([-1, 0, 5, 10, 11, 50, 100], h * w)
    quat = NS(x=0.0, y=0.0, z=math.sin(yaw / 2), w=math.cos(yaw / 2))
    msg = NS(info=NS(width=w, height=h, resolution=0.05,
                     origin=NS(position=NS(x=1.5, y=-2.0, z=0.0), orientation=quat)),
             data=[int(v) for v in data])
    m = rl.PyOMap(msg)
    grid = data.reshape(h, w)
    assert (m.width(), m.height()) == (h, w)
    for x in range(h):
        for y in range(w):
            assert m.isOccupied(x, y) == bool(grid[x, y] > 10), (x, y)
    scale, angle, ox, oy, s, c = m.world
    assert (scale, ox, oy) == (0.05, 1.5, -2.0)
    assert abs(angle + yaw) < 1e-12 and abs(s - math.sin(-yaw)) < 1e-12 and abs(c - math.cos(-yaw)) < 1e-12


# --------------------------------------------------------------------------- distance transform

@pytest.mark.parametrize("key", PNG_KEYS)
def test_edt_golden_png_maps(key):
    z = npz("png_maps.npz")
    img = unpack(z, f"{key}_t128")
    got = O.edt(img)
    if f"{key}_edt" in z:
        assert_bits_equal(got, z[f"{key}_edt"], key)
    else:
        assert_bits_equal(got.reshape(-1)[z[f"{key}_edt_idx"]], z[f"{key}_edt_val"], key)
        assert sha(got) == str(z[f"{key}_edt_sha"])


def test_edt_golden_generated_maps():
    for name, img in (("rect_disk_512.npz", W.rect_disk_map(512, seed=0)),
                      ("corridor_2048.npz", W.corridor_map(2048, seed=1))):
        z = npz(name)
        got = O.edt(img)
        assert sha(got) == str(z["edt_sha"]), name
        assert_bits_equal(got.reshape(-1)[z["edt_idx"]], z["edt_val"], name)


def test_edt_golden_degenerate():
    z = npz("degenerate.npz")
    for name in ("empty", "full", "one_free", "checker", "tall"):
        assert_bits_equal(O.edt(z[f"{name}_img"]), z[f"{name}_edt"], name)


def _edt_bruteforce(img):
    """Exact squared distance to the nearest occupied cell, then float sqrt;
    FLT_MAX where no cell is occupied (the reference's f = FLT_MAX start value)."""
    h, w = img.shape
    occ = np.argwhere(img.T != 0)  # (x, y)
    out = np.full((w, h), FLT_MAX, np.float32)
    if len(occ):
        xs, ys = np.meshgrid(np.arange(w), np.arange(h), indexing="ij")
        d2 = np.full((w, h), np.iinfo(np.int64).max, np.int64)
        for cx, cy in occ:
            d2 = np.minimum(d2, (xs - cx) ** 2 + (ys - cy) ** 2)
        out = d2.astype(np.float32)
    return np.sqrt(out).astype(np.float32)


@pytest.mark.parametrize("seed", range(12))
def test_edt_matches_bruteforce(seed):
    rng = np.random.default_rng(seed)
    h, w = int(rng.integers(1, 40)), int(rng.integers(1, 40))
    img = (rng.random((h, w)) < rng.choice([0.0, 0.01, 0.05, 0.3, 1.0])).astype(np.uint8)
    assert_bits_equal(O.edt(img), _edt_bruteforce(img), f"seed {seed}")


def test_edt_special_values():
    assert np.all(O.edt(np.zeros((5, 9), np.uint8)) == np.float32(np.sqrt(np.float32(FLT_MAX))))
    assert np.all(O.edt(np.ones((4, 3), np.uint8)) == 0)


# --------------------------------------------------------------------------- ranges vs golden vectors

def test_ranges_golden_png_maps():
    z = npz("png_maps.npz")
    for key in PNG_KEYS:
        img = unpack(z, f"{key}_t128")
        for kind in ("rm", "bl"):
            o = O.Oracle(img, 500.0, kind)
            eq = z[f"{key}_{kind}_edge_ins"]
            assert_bits_equal(o.calc_range_many(eq), z[f"{key}_{kind}_edge_grid"], f"{key} {kind} edge grid")
            assert_bits_equal(o.numpy_calc_range(eq), z[f"{key}_{kind}_edge_world"], f"{key} {kind} edge world")
            if f"{key}_{kind}_q" in z:
                assert_bits_equal(o.numpy_calc_range(z[f"{key}_{kind}_q"]), z[f"{key}_{kind}_world"], key)


@pytest.mark.parametrize("kind", ["rm", "bl"])
def test_ranges_golden_rect_disk_512(kind):
    z = npz("rect_disk_512.npz")
    img = W.rect_disk_map(512, seed=0)
    o = O.Oracle(img, 500.0, kind)
    o.set_sensor_model(W.sensor_table(501))
    assert_bits_equal(o.numpy_calc_range(z[f"{kind}_qw"]), z[f"{kind}_world"], "world")
    assert_bits_equal(o.calc_range_many(z[f"{kind}_qg"]), z[f"{kind}_grid"], "grid")
    ang = W.beam_angles(60)
    rr = o.numpy_calc_range_angles(z[f"{kind}_parts"], ang)
    assert_bits_equal(rr, z[f"{kind}_angles"], "angles")
    obs = z[f"{kind}_obs"]
    assert np.array_equal(o.eval_sensor_model(obs, rr, 60, 50), z[f"{kind}_sensor"])
    assert np.array_equal(o.calc_range_repeat_angles_eval_sensor_model(z[f"{kind}_parts"], ang, obs),
                          z[f"{kind}_fused"])
    ow = O.Oracle(img, 500.0, kind, tuple(z["world"]))
    assert_bits_equal(ow.numpy_calc_range(z[f"{kind}_wq"]), z[f"{kind}_wworld"], "world params")
    assert_bits_equal(ow.numpy_calc_range_angles(z[f"{kind}_wq"][:40], ang), z[f"{kind}_wangles"], "world angles")


@pytest.mark.parametrize("kind", ["rm", "bl"])
def test_ranges_golden_corridor_2048(kind):
    z = npz("corridor_2048.npz")
    img = W.corridor_map(2048, seed=1)
    assert_bits_equal(O.Oracle(img, 500.0, kind).numpy_calc_range_angles(z["parts"], W.beam_angles(60)),
                      z[f"{kind}_angles"], "angles")
    assert_bits_equal(O.Oracle(img, 1000.0, kind).numpy_calc_range(z[f"{kind}_qi"]), z[f"{kind}_r1000"], "r1000")


def test_ranges_golden_degenerate():
    z = npz("degenerate.npz")
    for name in ("empty", "full", "one_free", "checker", "tall"):
        img, eq = z[f"{name}_img"], z[f"{name}_ins"]
        for kind in ("rm", "bl"):
            for mr in (0, 1, 50):
                o = O.Oracle(img, float(mr), kind)
                assert_bits_equal(o.calc_range_many(eq), z[f"{name}_{kind}_{mr}_grid"], f"{name} {kind} {mr}")
                assert_bits_equal(o.numpy_calc_range(eq), z[f"{name}_{kind}_{mr}_world"], f"{name} {kind} {mr}")


# --------------------------------------------------------------------------- closed forms / brute force

def _bl_axis_scan(img, x, y, R, axis):
    """Bresenham along +x (heading 0) or +y (heading pi/2) from (x, y): the
    reference loop visits the cells at offsets k = 1.. while (int)(c + k) has
    not reached (int)(c + R + 1), returning k at the first occupied cell."""
    h, w = img.shape
    occ = img.T
    if 0 <= int(x) < w and 0 <= int(y) < h and occ[int(x), int(y)]:
        return 0.0
    c = np.float32(x if axis == 0 else y)
    end = int(np.float32(c + np.float32(R) + np.float32(1)))
    k, pos = 0, c
    while int(pos) != end:
        pos = np.float32(pos + np.float32(1))
        k += 1
        cx, cy = (int(pos), int(y)) if axis == 0 else (int(x), int(pos))
        if pos >= 0 and 0 <= cx < w and 0 <= cy < h and occ[cx, cy]:
            return float(k)
    return float(R)


@pytest.mark.parametrize("seed", range(6))
def test_bresenham_axis_rays_reduce_to_scans(seed):
    rng = np.random.default_rng(100 + seed)
    img = (rng.random((30, 30)) < 0.08).astype(np.uint8)
    R = float(rng.integers(1, 25))
    o = O.Oracle(img, R, "bl")
    for _ in range(80):
        x, y = float(rng.integers(0, 30)), float(rng.integers(0, 30))
        assert o.calc_range(x, y, 0.0) == _bl_axis_scan(img, x, y, R, 0)
        assert o.calc_range(x, y, np.float32(math.pi / 2)) == _bl_axis_scan(img, x, y, R, 1)


def test_rm_hit_is_an_occupied_cell_distance():
    """r < R implies r = sqrtf((cx - x0)^2 + (cy - y0)^2) for an occupied cell."""
    rng = np.random.default_rng(3)
    img = (rng.random((40, 40)) < 0.05).astype(np.uint8)
    o = O.Oracle(img, 60.0, "rm")
    occ = np.argwhere(img.T != 0).astype(np.float32)
    q = W.grid_poses(img, 300, seed=4)
    r = o.calc_range_many(q)
    assert np.all((r >= 0) & (r <= 60.0))
    for (x0, y0, _), ri in zip(q, r):
        if ri < 60.0:
            dx, dy = occ[:, 0] - np.float32(x0), occ[:, 1] - np.float32(y0)
            assert np.any(np.sqrt(dx * dx + dy * dy) == ri)


@pytest.mark.parametrize("kind", ["rm", "bl"])
def test_world_transform_identity_closed_form(kind):
    """With identity world parameters numpy_calc_range(x, y, th) casts from grid
    (y, x) at heading -th + fl(-3*pi/2) (RangeLib.h:446-475)."""
    img = W.rect_disk_map(128, seed=5)
    o = O.Oracle(img, 100.0, kind)
    q = W.world_poses(img, 500, seed=6)
    rot = np.float32(-1.0 * 0.0 - 3.0 * math.pi / 2.0)
    g = np.stack([q[:, 1], q[:, 0], (-q[:, 2]) + rot], 1).astype(np.float32)
    assert_bits_equal(o.numpy_calc_range(q), o.calc_range_many(g), kind)


def test_sensor_model_power_of_two_table():
    """table[r][d] = 2^-(r + 2d): the weight is the exact power of two
    2^-sum(r_j + 2 d_ij), which fixes row = observed, column = expected range,
    the clamping to [0, K-1] and truncation of both indices."""
    K = 20
    r, d = np.meshgrid(np.arange(K), np.arange(K), indexing="ij")
    table = np.ldexp(1.0, -(r + 2 * d)).astype(np.float64)
    o = O.Oracle(np.zeros((4, 4), np.uint8), 10.0)
    o.set_sensor_model(table)
    rng = np.random.default_rng(8)
    M, N = 13, 40
    obs = rng.uniform(-3, 25, M).astype(np.float32)
    rngs = rng.uniform(-3, 25, N * M).astype(np.float32)
    w = o.eval_sensor_model(obs, rngs, M, N)
    ri = np.clip(obs, 0, K - 1).astype(np.int64)
    di = np.clip(rngs.reshape(N, M), 0, K - 1).astype(np.int64)
    want = np.ldexp(1.0, -(ri[None, :] + 2 * di).sum(1))
    assert np.array_equal(w, want)


def test_sensor_model_constant_table():
    c = 0.37
    o = O.Oracle(np.zeros((4, 4), np.uint8), 10.0)
    o.set_sensor_model(np.full((11, 11), c))
    w = o.eval_sensor_model(np.ones(7, np.float32), np.ones(7 * 3, np.float32), 7, 3)
    want = 1.0
    for _ in range(7):
        want *= c
    assert np.all(w == want)


# --------------------------------------------------------------------------- trig

def test_trig_table_exhaustive_vs_libc():
    """Every float with |x| <= 64: the GPU's sinf/cosf algorithm (run on the host)
    equals the C library bit for bit."""
    from range_libc_amd import _lib

    exe = os.path.join(_lib.PKG, "_data", "gen_trig_table")
    table = os.path.join(_lib.PKG, "_data", "trig_exc.bin")
    p = subprocess.run([exe, "--verify", table], capture_output=True, text=True, timeout=300)
    assert p.returncode == 0, p.stdout + p.stderr
    assert " 0 mismatches" in p.stdout


# --------------------------------------------------------------------------- live vs the compiled reference

needs_ref = pytest.mark.skipif(not O.ref_available(), reason="compiled reference not built here")


@needs_ref
@pytest.mark.parametrize("seed", range(8))
def test_oracle_matches_compiled_reference(seed):
    rng = np.random.default_rng(1000 + seed)
    h, w = int(rng.integers(2, 200)), int(rng.integers(2, 200))
    img = (rng.random((h, w)) < rng.uniform(0.001, 0.2)).astype(np.uint8)
    assert_bits_equal(O.edt(img), O.ref_edt(img), "edt")
    world = None
    if seed % 2:
        a = float(rng.uniform(-3, 3))
        world = (float(rng.uniform(0.02, 2)), a, float(rng.uniform(-5, 5)), float(rng.uniform(-5, 5)),
                 math.sin(a), math.cos(a))
    R = float(rng.uniform(1, 300))
    q = np.stack([rng.uniform(-10, w + 10, 2000), rng.uniform(-10, h + 10, 2000),
                  rng.uniform(-20, 20, 2000)], 1).astype(np.float32)
    table = rng.random((31, 31))
    for kind in ("rm", "bl"):
        o = O.Oracle(img, R, kind, world or O.IDENTITY_WORLD)
        r = O.Reference(img, R, kind, world)
        o.set_sensor_model(table)
        r.set_sensor_model(table)
        assert_bits_equal(o.calc_range_many(q), r.calc_range_many(q), "grid")
        assert_bits_equal(o.numpy_calc_range(q), r.numpy_calc_range(q), "world")
        ang = rng.uniform(-3, 3, 17).astype(np.float32)
        rr = o.numpy_calc_range_angles(q[:50], ang)
        assert_bits_equal(rr, r.numpy_calc_range_angles(q[:50], ang), "angles")
        obs = rng.uniform(-2, 40, 17).astype(np.float32)
        assert np.array_equal(o.eval_sensor_model(obs, rr, 17, 50), r.eval_sensor_model(obs, rr, 17, 50))
        assert np.array_equal(o.calc_range_repeat_angles_eval_sensor_model(q[:50], ang, obs),
                              r.calc_range_repeat_angles_eval_sensor_model(q[:50], ang, obs))


# --------------------------------------------------------------------------- GiantLUTCast

def _glt_fixture():
    return np.load(os.path.join(GOLDEN, "giant_lut.npz"))


def _glt_names(g):
    return sorted({k.rsplit("_img", 1)[0] for k in g.files if k.endswith("_img")})


def test_glt_oracle_golden():
    """Table (hash + samples), constants and queries of the compiled reference's
    GiantLUTCast equal the oracle's (grid, world, rotated world, repeated angles)."""
    g = _glt_fixture()
    world = (0.05, 0.3, -3.0, 2.0, math.sin(0.3), math.cos(0.3))
    for name in _glt_names(g):
        img = g[f"{name}_img"]
        R = {"rect64": 60.0, "corr96x80": 120.0, "rect48": 500.0}[name]
        td = {"rect64": 108, "corr96x80": 37, "rect48": 360}[name]
        o = O.Oracle(img, R, "glt", theta_disc=td)
        assert hashlib.sha256(np.ascontiguousarray(o.glt).tobytes()).hexdigest() == str(g[f"{name}_table_sha"])
        assert np.array_equal(o.glt.reshape(-1)[g[f"{name}_sample_idx"]], g[f"{name}_sample"])
        assert o.m.glt_max_div_limits == float(g[f"{name}_consts"][0])
        q = g[f"{name}_ins"]
        assert_bits_equal(o.calc_range_many(q), g[f"{name}_grid"], f"{name} grid")
        assert_bits_equal(o.numpy_calc_range(q), g[f"{name}_world"], f"{name} world")
        assert_bits_equal(o.numpy_calc_range_angles(q[:300], W.beam_angles(61)), g[f"{name}_angles"], f"{name} angles")
        ow = O.Oracle(img, R, "glt", world, theta_disc=td)
        assert_bits_equal(ow.numpy_calc_range(q), g[f"{name}_world_rot"], f"{name} rotated world")


def test_glt_bins_and_entries_closed_form():
    """A heading i * (2 pi / td) (the build's own angles) falls in bin i; a heading
    shifted by +-2 pi wraps back to the same bin when the wrap is exact in float;
    every table entry is the quantized RayMarching range from the cell corner."""
    td_list = (1, 7, 37, 108, 360, 1000)
    for td in td_list:
        step = np.float32(6.28318530718 / np.float64(np.float32(td)))
        c = np.float32(np.float64(td) / 6.28318530718)
        for i in range(td):
            assert O.glt_discretize(float(np.float32(i) * step), td, float(c)) == i
    assert O.glt_discretize(float("nan"), 108, float(np.float32(108 / 6.28318530718))) == -1
    assert O.glt_discretize(1e10, 108, float(np.float32(108 / 6.28318530718))) == -1
    img = W.rect_disk_map(40, seed=9)
    R, td = 50.0, 24
    o = O.Oracle(img, R, "glt", theta_disc=td)
    rm = O.Oracle(img, R, "rm")
    step = np.float32(6.28318530718 / np.float64(np.float32(td)))
    ldm = np.float32(65535) / np.float32(R)
    rng = np.random.default_rng(0)
    for _ in range(300):
        x, y, i = int(rng.integers(0, 40)), int(rng.integers(0, 40)), int(rng.integers(0, td))
        r = np.float32(rm.calc_range(float(x), float(y), float(np.float32(i) * step)))
        r = min(r, np.float32(R))
        assert o.glt[x, y, i] == np.uint16(np.float32(r * ldm))


@needs_ref
@pytest.mark.parametrize("seed", range(3))
def test_glt_oracle_matches_compiled_reference(seed):
    rng = np.random.default_rng(2000 + seed)
    h, w = int(rng.integers(2, 50)), int(rng.integers(2, 50))
    img = (rng.random((h, w)) < rng.uniform(0.01, 0.3)).astype(np.uint8)
    R, td = float(rng.uniform(5, 200)), int(rng.integers(1, 200))
    o = O.Oracle(img, R, "glt", theta_disc=td)
    r = O.Reference(img, R, "glt", theta_disc=td)
    table, consts = r.glt_table()
    assert np.array_equal(o.glt, table)
    q = np.stack([rng.uniform(-3, w + 3, 3000), rng.uniform(-3, h + 3, 3000),
                  rng.uniform(-100, 100, 3000)], 1).astype(np.float32)
    assert_bits_equal(o.calc_range_many(q), r.calc_range_many(q), "grid")
    assert_bits_equal(o.numpy_calc_range(q), r.numpy_calc_range(q), "world")


# --------------------------------------------------------------------------- CDDTCast / PCDDT

CDDT_CASES = {"rect64": (60.0, 108), "corr96x80": (120.0, 37), "corr300": (200.0, 120)}
CDDT_WORLD = (0.05, 0.3, -3.0, 2.0, math.sin(0.3), math.cos(0.3))


def _sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


@pytest.mark.parametrize("name", sorted(CDDT_CASES))
def test_cddt_oracle_golden(name):
    """Tables (full and pruned) and queries of the compiled reference's CDDTCast
    equal the oracle's."""
    from tests.golden.make_golden import world_inside

    g = np.load(os.path.join(GOLDEN, "cddt.npz"))
    R, td = CDDT_CASES[name]
    img = g[f"{name}_img"]
    q = g[f"{name}_ins"]
    for tag, prune in (("full", False), ("pruned", True)):
        o = O.Oracle(img, R, "cddt", theta_disc=td, prune=prune)
        t = o.cddt_tables()
        assert t["widths"].astype(np.int32).tobytes() and _sha(t["widths"].astype(np.int32)) == str(g[f"{name}_{tag}_widths_sha"])
        assert _sha(t["trans"]) == str(g[f"{name}_{tag}_trans_sha"])
        assert _sha(t["start"].astype(np.int64)) == str(g[f"{name}_{tag}_start_sha"])
        assert _sha(t["vals"]) == str(g[f"{name}_{tag}_vals_sha"])
        assert_bits_equal(o.calc_range_many(q), g[f"{name}_{tag}_grid"], f"{name} {tag} grid")
    h, w = img.shape
    ow = O.Oracle(img, R, "cddt", CDDT_WORLD, theta_disc=td)
    assert_bits_equal(ow.numpy_calc_range(world_inside(q, CDDT_WORLD, w, h)), g[f"{name}_world_rot"],
                      f"{name} rotated world")


def test_cddt_pruning_keeps_every_answer():
    """Pruning removes table entries but no query from a free cell centre changes
    (the entries it drops are never returned), and the pruned table is smaller."""
    img = W.corridor_map(160, seed=8)
    R, td = 150.0, 64
    full = O.Oracle(img, R, "cddt", theta_disc=td)
    pruned = O.Oracle(img, R, "cddt", theta_disc=td, prune=True)
    assert 0 < len(pruned.cddt_tables()["vals"]) < len(full.cddt_tables()["vals"])
    cells = W.free_cells(img)[::7].astype(np.float32) + 0.5
    step = np.float32(2 * math.pi / td)
    for i in range(0, td, 5):
        q = np.concatenate([cells, np.full((len(cells), 1), np.float32(i) * step)], axis=1).astype(np.float32)
        a, b = full.calc_range_many(q), pruned.calc_range_many(q)
        ok = a < np.float32(R)  # a miss beyond max_range may hit a dropped entry; those return max_range unpruned
        assert_bits_equal(a[ok], b[ok], f"bin {i}")


@needs_ref
@pytest.mark.parametrize("seed", range(3))
def test_cddt_oracle_matches_compiled_reference(seed):
    rng = np.random.default_rng(3000 + seed)
    h, w = int(rng.integers(8, 70)), int(rng.integers(8, 70))
    img = (rng.random((h, w)) < rng.uniform(0.02, 0.3)).astype(np.uint8)
    img[0, :] = img[-1, :] = 1
    img[:, 0] = img[:, -1] = 1
    R, td = float(rng.uniform(5, 200)), int(rng.integers(2, 200))
    for prune in (False, True):
        o = O.Oracle(img, R, "cddt", theta_disc=td, prune=prune)
        r = O.Reference(img, R, "cddt", theta_disc=td)
        if prune:
            r.cddt_prune(R)
        a, b = o.cddt_tables(), r.cddt_tables()
        for k in ("widths", "start"):
            assert np.array_equal(a[k], b[k]), k
        assert_bits_equal(a["trans"], b["trans"], "trans")
        assert_bits_equal(a["vals"], b["vals"], "vals")
        q = np.stack([rng.uniform(0, w, 3000), rng.uniform(0, h, 3000), rng.uniform(-100, 100, 3000)],
                     1).astype(np.float32)
        assert_bits_equal(o.calc_range_many(q), r.calc_range_many(q), "grid")


@pytest.mark.parametrize("name", sorted(CDDT_CASES))
def test_cddt_radial_optimized_golden(name):
    g = np.load(os.path.join(GOLDEN, "cddt.npz"))
    R, td = CDDT_CASES[name]
    img = g[f"{name}_img"]
    ins = W.world_poses(img, 200, seed=6)
    for tag, prune in (("full", False), ("pruned", True)):
        o = O.Oracle(img, R, "cddt", theta_disc=td, prune=prune)
        got = o.calc_range_many_radial_optimized(ins, 60, -3 * math.pi / 4, 3 * math.pi / 4)
        assert_bits_equal(got, g[f"{name}_{tag}_radial"], f"{name} {tag} radial")


@needs_ref
@pytest.mark.parametrize("nr,mn,mx", [(60, -3 * math.pi / 4, 3 * math.pi / 4), (100, -math.pi, math.pi),
                                      (31, -math.pi / 2, math.pi / 2 + 0.3)])
def test_radial_optimized_matches_compiled_reference(nr, mn, mx):
    img = W.corridor_map(200, seed=3)
    ins = W.world_poses(img, 150, seed=nr)
    for kind, td, prune in (("cddt", 108, False), ("cddt", 37, True), ("rm", 0, False)):
        ref = O.Reference(img, 150.0, kind, theta_disc=td)
        if prune:
            ref.cddt_prune(150.0)
        o = O.Oracle(img, 150.0, kind, theta_disc=td, prune=prune)
        assert_bits_equal(o.calc_range_many_radial_optimized(ins, nr, mn, mx),
                          ref.calc_range_many_radial_optimized(ins, nr, mn, mx), f"{kind} {td} {prune}")


@pytest.mark.parametrize("name", sorted(CDDT_CASES))
def test_cddt_json_golden(name):
    """cddt_json over the oracle's tables reproduces the reference's serializeJson text."""
    import range_libc_amd as rl

    g = np.load(os.path.join(GOLDEN, "cddt.npz"))
    R, td = CDDT_CASES[name]
    img = g[f"{name}_img"]
    for tag, prune in (("full", False), ("pruned", True)):
        t = O.Oracle(img, R, "cddt", theta_disc=td, prune=prune).cddt_tables()
        text = rl.cddt_json(td, t["trans"], R, "", O.xmajor(img), t["widths"], t["start"], t["vals"])
        assert hashlib.sha256(text.encode()).hexdigest() == str(g[f"{name}_{tag}_json_sha"])


def test_sensor_model_nan_index_is_max_row_and_column():
    """A NaN observation or expected range selects the table's last row /
    column (the reference indexes outside its table there)."""
    img = np.zeros((16, 16), np.uint8)
    K = 9
    table = (np.arange(K * K, dtype=np.float64).reshape(K, K) + 1.0) / 64.0
    orc = O.Oracle(img, 8.0, "rm")
    orc.set_sensor_model(table)
    obs = np.array([np.nan, 2.0, 3.5], np.float32)
    ranges = np.array([1.0, np.nan, 7.9, np.nan, np.nan, 100.0], np.float32)
    w = orc.eval_sensor_model(obs, ranges, 3, 2)
    want0 = table[K - 1, 1] * table[2, K - 1] * table[3, 7]
    want1 = table[K - 1, K - 1] * table[2, K - 1] * table[3, K - 1]
    assert w[0] == ((1.0 * table[K - 1, 1]) * table[2, K - 1]) * table[3, 7] and np.isclose(w[0], want0)
    assert w[1] == ((1.0 * table[K - 1, K - 1]) * table[2, K - 1]) * table[3, K - 1] and np.isclose(w[1], want1)
