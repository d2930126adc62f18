"""GPU parity: the HIP library against the CPU oracle, bit for bit.

Every test runs the product path (librangelib_hip.so through its C ABI) on
the device and compares with the oracle restatement of the reference's CPU
code on the same seeded inputs.
"""
import math

import numpy as np
import pytest

import range_libc_amd as rl
from oracle import oracle as O
from range_libc_amd import workloads as W
from tests.conftest import assert_bits_equal

pytestmark = pytest.mark.gpu


def _torch():
    import torch

    assert torch.cuda.is_available(), "GPU test needs a device"
    return torch


# --------------------------------------------------------------------------- trig / sqrt

def test_sincosf_bit_exact_vs_libc(ctx):
    torch = _torch()
    rng = np.random.default_rng(0)
    two_to_four = np.arange(0x40000000, 0x40800000, dtype=np.uint32).view(np.float32)  # all of [2, 4)
    xs = np.concatenate([
        rng.uniform(-64, 64, 3_000_000).astype(np.float32),
        rng.uniform(-16, 0, 1_000_000).astype(np.float32),
        two_to_four, -two_to_four,
        np.array([0.0, -0.0, 1e-30, -1e-30, 64.0, -64.0, math.pi, -math.pi], dtype=np.float32),
        rng.uniform(-1e4, 1e4, 200_000).astype(np.float32),  # outside the table domain
    ])
    x = torch.from_numpy(xs).cuda()
    s = torch.empty_like(x)
    c = torch.empty_like(x)
    ctx.sincosf_dev(x.data_ptr(), s.data_ptr(), c.data_ptr(), x.numel())
    ctx.sync()
    s_h, c_h = s.cpu().numpy(), c.cpu().numpy()
    inside = np.abs(xs) <= 64.0
    assert_bits_equal(s_h[inside], O.libc_sinf(xs[inside]), "sinf |x|<=64")
    assert_bits_equal(c_h[inside], O.libc_cosf(xs[inside]), "cosf |x|<=64")
    # outside the exhaustively tabulated domain the result is the nearest float;
    # the C library misrounds a few percent of such large arguments (1.3% here)
    out = ~inside
    ds = np.abs(s_h[out].view(np.int32).astype(np.int64) - O.libc_sinf(xs[out]).view(np.int32))
    assert ds.max() <= 1 and (ds > 0).mean() < 0.03


def test_sqrtf_correctly_rounded(ctx):
    torch = _torch()
    rng = np.random.default_rng(1)
    xs = np.concatenate([np.arange(0, 1 << 24, dtype=np.float32),
                         rng.uniform(0, 1e7, 2_000_000).astype(np.float32),
                         np.array([np.finfo(np.float32).max, 0.0, 1e-40], dtype=np.float32)])
    x = torch.from_numpy(xs).cuda()
    o = torch.empty_like(x)
    ctx.sqrtf_dev(x.data_ptr(), o.data_ptr(), x.numel())
    ctx.sync()
    assert_bits_equal(o.cpu().numpy(), np.sqrt(xs), "sqrtf")


# --------------------------------------------------------------------------- distance transform

def _points(shape, pts):
    img = np.zeros(shape, np.uint8)
    for y, x in pts:
        img[y, x] = 1
    return img


def _edt_gpu(ctx, img, flags=0):
    m = rl.DeviceMap(ctx, O.xmajor(img), flags)
    return m.edt()


@pytest.mark.parametrize("case", ["rect512", "corridor1024", "empty", "full", "single_free",
                                  "one_px", "tall", "wide", "sparse_cols", "very_tall", "mixed_rows",
                                  "wider_than_4096", "width_4092", "width_4096"])
def test_edt_bit_exact(ctx, case):
    rng = np.random.default_rng(7)
    img = {
        "rect512": lambda: W.rect_disk_map(512, seed=2),
        "corridor1024": lambda: W.corridor_map(1024, seed=3),
        "empty": lambda: np.zeros((37, 53), np.uint8),
        "full": lambda: np.ones((20, 30), np.uint8),
        "single_free": lambda: np.pad(np.zeros((1, 1), np.uint8), 3, constant_values=1),
        "one_px": lambda: np.ones((1, 1), np.uint8),
        "tall": lambda: (rng.random((700, 13)) < 0.02).astype(np.uint8),
        "wide": lambda: (rng.random((9, 800)) < 0.01).astype(np.uint8),
        "sparse_cols": lambda: (rng.random((300, 257)) < 0.0005).astype(np.uint8),
        "very_tall": lambda: (rng.random((4500, 6)) < 0.003).astype(np.uint8),  # > 4096 rows
        # rows y >= 926 have g + x^2 >= 2^24 at x = 3990: the row search hands them to the envelope
        "mixed_rows": lambda: _points((1500, 4000), [(0, 3990), (700, 10)]),
        "wider_than_4096": lambda: (rng.random((5, 4200)) < 0.002).astype(np.uint8),  # envelope on every row
        # the widest rows the LDS row search takes, and just past it (the envelope)
        "width_4092": lambda: (rng.random((30, 4092)) < 0.003).astype(np.uint8),
        "width_4096": lambda: (rng.random((30, 4096)) < 0.003).astype(np.uint8),
    }[case]()
    want = O.edt(img)
    assert_bits_equal(_edt_gpu(ctx, img), want, f"edt {case}")
    # the full lower-envelope path on columns (used above 4096 rows) agrees too
    assert_bits_equal(_edt_gpu(ctx, img, rl._lib.RL_MAP_EDT_ENVELOPE_ONLY), want, f"edt envelope {case}")


# --------------------------------------------------------------------------- ray casting

@pytest.fixture(scope="module")
def rect_map():
    return W.rect_disk_map(512, seed=0)


@pytest.fixture(scope="module")
def corridor_map():
    return W.corridor_map(2048, seed=1)


@pytest.mark.parametrize("kind", ["rm", "bl"])
def test_calc_range_many_world_frame(kind, rect_map):
    ins = W.world_poses(rect_map, 100_000, seed=5)
    cls = rl.PyRayMarchingGPU if kind == "rm" else rl.PyBresenhamsLine
    meth = cls(rl.PyOMap(rect_map), 500.0)
    out = np.empty(ins.shape[0], np.float32)
    meth.calc_range_many(ins, out)
    assert_bits_equal(out, O.Oracle(rect_map, 500.0, kind).numpy_calc_range(ins), f"{kind} world")


@pytest.mark.parametrize("kind", ["rm", "bl"])
def test_calc_range_many_grid_frame(kind, rect_map):
    ins = W.grid_poses(rect_map, 50_000, seed=6)
    cls = rl.PyRayMarchingGPU if kind == "rm" else rl.PyBresenhamsLine
    meth = cls(rl.PyOMap(rect_map), 500.0)
    out = np.empty(ins.shape[0], np.float32)
    meth.calc_range_many_grid(ins, out)
    assert_bits_equal(out, O.Oracle(rect_map, 500.0, kind).calc_range_many(ins), f"{kind} grid")


@pytest.mark.parametrize("kind", ["rm", "bl"])
def test_repeat_angles(kind, corridor_map):
    ins = W.world_poses(corridor_map, 4000, seed=7)
    ang = W.beam_angles(60)
    cls = rl.PyRayMarchingGPU if kind == "rm" else rl.PyBresenhamsLine
    meth = cls(rl.PyOMap(corridor_map), 500.0)
    out = np.empty(4000 * 60, np.float32)
    meth.calc_range_repeat_angles(ins, ang, out)
    assert_bits_equal(out, O.Oracle(corridor_map, 500.0, kind).numpy_calc_range_angles(ins, ang),
                      f"{kind} angles")


def test_world_params_non_identity(rect_map):
    world = (0.05, 0.3, -3.0, 2.0, math.sin(0.3), math.cos(0.3))
    omap = rl.PyOMap(rect_map)
    omap.set_world(*world)
    # poses whose grid image lies in free space: invert x = R(angle) (x_w - o) / scale
    g = W.world_poses(rect_map, 20_000, seed=8).astype(np.float64)
    c, s = math.cos(0.3), math.sin(0.3)
    xw = (c * g[:, 0] + s * g[:, 1]) * 0.05 - 3.0
    yw = (-s * g[:, 0] + c * g[:, 1]) * 0.05 + 2.0
    wins = np.stack([xw, yw, g[:, 2] - 1.0], axis=1).astype(np.float32)
    for kind, cls in (("rm", rl.PyRayMarchingGPU), ("bl", rl.PyBresenhamsLine)):
        out = np.empty(wins.shape[0], np.float32)
        cls(omap, 500.0).calc_range_many(wins, out)
        assert_bits_equal(out, O.Oracle(rect_map, 500.0, kind, world).numpy_calc_range(wins), f"{kind} world params")


def test_single_query_calc_range(rect_map):
    orc = O.Oracle(rect_map, 300.0, "rm")
    orb = O.Oracle(rect_map, 300.0, "bl")
    rm = rl.PyRayMarching(rl.PyOMap(rect_map), 300.0)
    bl = rl.PyBresenhamsLine(rl.PyOMap(rect_map), 300.0)
    for x, y, h in W.grid_poses(rect_map, 20, seed=9):
        assert np.float32(rm.calc_range(x, y, h)) == np.float32(orc.calc_range(x, y, h))
        assert np.float32(bl.calc_range(x, y, h)) == np.float32(orb.calc_range(x, y, h))


# --------------------------------------------------------------------------- sensor model

@pytest.mark.parametrize("m", [1, 7, 60, 64, 65, 120])
def test_eval_sensor_model_bit_exact(m, corridor_map):
    rng = np.random.default_rng(m)
    table = W.sensor_table(501)
    n = 1000
    ranges = rng.uniform(-5, 520, n * m).astype(np.float32)
    obs = rng.uniform(-5, 520, m).astype(np.float32)
    meth = rl.PyRayMarchingGPU(rl.PyOMap(np.zeros((8, 8), np.uint8)), 500.0)
    meth.set_sensor_model(table)
    w = np.empty(n, np.float64)
    meth.eval_sensor_model(obs, ranges, w, m, n)
    orc = O.Oracle(np.zeros((8, 8), np.uint8), 500.0)
    orc.set_sensor_model(table)
    assert np.array_equal(w, orc.eval_sensor_model(obs, ranges, m, n))


@pytest.mark.parametrize("kind,m", [("rm", 60), ("rm", 30), ("rm", 100), ("rm", 64), ("rm", 65), ("rm", 128),
                                    ("bl", 60)])
def test_fused_sensor_model_bit_exact(kind, m, corridor_map):
    table = W.sensor_table(501)
    ins = W.world_poses(corridor_map, 2000, seed=10)
    ang = W.beam_angles(m)
    orc = O.Oracle(corridor_map, 500.0, kind)
    orc.set_sensor_model(table)
    true_pose = W.world_poses(corridor_map, 1, seed=11)
    obs = W.noisy_observation(orc.numpy_calc_range_angles(true_pose, ang), seed=12)
    cls = rl.PyRayMarchingGPU if kind == "rm" else rl.PyBresenhamsLine
    meth = cls(rl.PyOMap(corridor_map), 500.0)
    meth.set_sensor_model(table)
    w = np.empty(2000, np.float64)
    meth.calc_range_repeat_angles_eval_sensor_model(ins, ang, obs, w)
    assert np.array_equal(w, orc.calc_range_repeat_angles_eval_sensor_model(ins, ang, obs))


@pytest.mark.parametrize("n,m", [(50, 60), (3000, 60), (50_000, 8)])
def test_host_calls_zero_copy_and_staged(n, m, corridor_map):
    """The host-pointer sensor-model and fused calls run zero-copy through a
    pinned block when their arrays fit 512 KiB (50 x 60) and through device
    staging otherwise (3000 x 60 ranges: 720 KB; 50,000 poses: 600 KB);
    both give the oracle's weights."""
    table = W.sensor_table(501)
    R = 300.0
    ins = W.world_poses(corridor_map, n, seed=40 + m)
    ang = W.beam_angles(m)
    orc = O.Oracle(corridor_map, R, "rm")
    orc.set_sensor_model(table)
    obs = W.noisy_observation(orc.numpy_calc_range_angles(W.world_poses(corridor_map, 1, seed=41), ang), seed=42,
                              max_range=R)
    meth = rl.PyRayMarchingGPU(rl.PyOMap(corridor_map), R)
    meth.set_sensor_model(table)
    ranges = np.empty(n * m, np.float32)
    meth.calc_range_repeat_angles(ins, ang, ranges)
    assert_bits_equal(ranges, orc.numpy_calc_range_angles(ins, ang), "ranges")
    w = np.empty(n, np.float64)
    meth.eval_sensor_model(obs, ranges, w, m, n)
    assert np.array_equal(w, orc.eval_sensor_model(obs, ranges, m, n))
    wf = np.empty(n, np.float64)
    meth.calc_range_repeat_angles_eval_sensor_model(ins, ang, obs, wf)
    assert np.array_equal(wf, orc.calc_range_repeat_angles_eval_sensor_model(ins, ang, obs))


def test_sensor_model_requires_table():
    meth = rl.PyRayMarchingGPU(rl.PyOMap(np.zeros((8, 8), np.uint8)), 10.0)
    with pytest.raises(rl.RangeLibError):
        meth.eval_sensor_model(np.zeros(4, np.float32), np.zeros(8, np.float32), np.zeros(2), 4, 2)


def test_empty_batches(rect_map):
    meth = rl.PyRayMarchingGPU(rl.PyOMap(rect_map), 100.0)
    meth.calc_range_many(np.zeros((0, 3), np.float32), np.zeros(0, np.float32))
    meth.calc_range_repeat_angles(np.zeros((0, 3), np.float32), np.zeros(4, np.float32), np.zeros(0, np.float32))


@pytest.mark.parametrize("flags", [0, "no_compact"])
def test_rm_special_headings_and_origins(ctx, rect_map, flags):
    """NaN and infinite headings march nowhere (the reference's float-to-int
    conversion of a NaN position is INT_MIN, outside the map: max_range); huge,
    signed-zero and wrapped headings, origins just below 0, on the border, NaN and
    outside the map, and max_range <= 0, all bit-exact with read counts."""
    flags = rl._lib.RL_MAP_NO_COMPACT if flags == "no_compact" else 0
    m = rl.DeviceMap(ctx, O.xmajor(rect_map), flags)
    # bit parity of the directions holds for |heading| <= 64 (sinf/cosf equal the C
    # library's there); the large headings below were checked to round like the C
    # library as well, so their rays stay comparable bit for bit
    heads = np.array([np.nan, np.inf, -np.inf, 1e6, -777777.0, 3e4, -3e4, 2 * np.pi, -0.0, 0.0, 1e-30, 4.0],
                     np.float32)
    origins = np.array([[100.5, 200.25], [-0.5, 30.0], [30.0, -0.75], [0.0, 0.0], [511.9, 511.9], [512.0, 3.0],
                        [np.nan, 5.0], [5.0, np.inf], [-3.0, 40.0], [250.0, 250.0]], np.float32)
    q = np.array([[x, y, h] for x, y in origins for h in heads], np.float32)
    n = q.shape[0]
    for R in (500.0, 0.0, -1.0):
        out = np.empty(n, np.float32)
        m.calc_range_many(rl.RL_METHOD_RM, R, q, out, n)
        want, want_cnt = O.Oracle(rect_map, R, "rm").calc_range_many(q, counts=True)
        assert_bits_equal(out, want, f"R={R}")
    torch = _torch()
    d_q = torch.from_numpy(q).cuda()
    cnt = torch.zeros(n, dtype=torch.int32, device="cuda")
    m.count_steps_dev(rl.RL_METHOD_RM, rl.RL_FRAME_GRID, 500.0, d_q.data_ptr(), 0, cnt.data_ptr(), n, 1)
    ctx.sync()
    _, want_cnt = O.Oracle(rect_map, 500.0, "rm").calc_range_many(q, counts=True)
    assert np.array_equal(cnt.cpu().numpy(), want_cnt)


def test_bresenham_vs_rm_full_size(corridor_map):
    """The Bresenham-vs-ray-marching workload at full size: 1 M queries from
    integer free cells of the 2048^2 corridor map, max_range 1000.  Every ray is
    cast on the GPU; 20,000 of them, drawn at random, are compared bit for bit
    with the oracle, and all of them satisfy 0 <= r <= R + 2 (a hit cell's
    distance can exceed R by less than the cell diagonal)."""
    torch = _torch()
    ctx = rl.default_context(0)
    n, R = 1_000_000, 1000.0
    q = W.world_poses(corridor_map, n, seed=7, integer=True)
    m = rl.DeviceMap(ctx, O.xmajor(corridor_map))
    d_q = torch.from_numpy(q).cuda()
    pick = np.random.default_rng(3).choice(n, 20_000, replace=False)
    for method, kind in ((rl.RL_METHOD_BL, "bl"), (rl.RL_METHOD_RM, "rm")):
        out = torch.full((n,), -1.0, dtype=torch.float32, device="cuda")
        m.numpy_calc_range_dev(method, R, d_q.data_ptr(), out.data_ptr(), n)
        ctx.sync()
        got = out.cpu().numpy()
        assert np.all(np.isfinite(got)) and got.min() >= 0.0 and got.max() <= R + 2.0, kind
        assert_bits_equal(got[pick], O.Oracle(corridor_map, R, kind).numpy_calc_range(q[pick]), f"{kind} sample")


@pytest.mark.parametrize("R", [1000.0, 40.0, 31.0])
def test_bresenham_all_rays(corridor_map, R):
    """Every one of 100,000 Bresenham rays on the 2048^2 corridor map equals the
    oracle, in the grid and the world frame, for ranges whose walks end within
    about 32 steps or run long, including origins off the map."""
    torch = _torch()
    ctx = rl.default_context(0)
    n = 100_000
    q = W.world_poses(corridor_map, n, seed=21)
    rng = np.random.default_rng(22)
    q[:500, 0] = rng.uniform(-60.0, 2108.0, 500)  # some origins off the map
    q[:500, 1] = rng.uniform(-60.0, 2108.0, 500)
    m = rl.DeviceMap(ctx, O.xmajor(corridor_map))
    d_q = torch.from_numpy(q).cuda()
    orc = O.Oracle(corridor_map, R, "bl")
    for name, call, want in (("world", m.numpy_calc_range_dev, orc.numpy_calc_range(q)),
                             ("grid", m.calc_range_many_dev, orc.calc_range_many(q))):
        out = torch.full((n,), -1.0, dtype=torch.float32, device="cuda")
        call(rl.RL_METHOD_BL, R, d_q.data_ptr(), out.data_ptr(), n)
        ctx.sync()
        assert_bits_equal(out.cpu().numpy(), want, f"{name} R={R}")


def test_bresenham_far_origins(ctx, rect_map):
    """Bresenham from origins far off the map along either axis (up to 3e9,
    beyond int range: the reference's loop then does not run, or walks exact
    integer steps below 2^24), NaN origins (no walk: max_range), on the border
    and just inside, with headings along and across the axes: every range equals
    the oracle's, which walks like the reference (none of these inputs makes the
    reference loop forever).  The walk converts the minor coordinate to a cell
    only after clamping it, so no conversion leaves int range."""
    m = rl.DeviceMap(ctx, O.xmajor(rect_map))
    heads = np.array([0.0, 0.3, 1.2, np.pi / 2, 2.0, np.pi, -0.7, -np.pi / 2, 4.0], np.float32)
    origins = np.array([[3e9, 50.0], [50.0, 3e9], [-3e9, 200.0], [200.0, -3e9], [1e7, 100.0], [100.0, -1e7],
                        [5e6, 5e6], [-40.0, 100.0], [100.0, 600.0], [511.0, 0.0], [0.0, 511.0], [250.5, 250.5],
                        [np.nan, 100.0], [100.0, np.nan]], np.float32)
    q = np.array([[x, y, h] for x, y in origins for h in heads], np.float32)
    n = q.shape[0]
    for R in (1000.0, 50.0):
        out = np.empty(n, np.float32)
        m.calc_range_many(rl.RL_METHOD_BL, R, q, out, n)
        assert_bits_equal(out, O.Oracle(rect_map, R, "bl").calc_range_many(q), f"R={R}")


def test_bresenham_rounding_steps(ctx):
    """Bresenham walks whose unit steps round: fractional origins just below a
    power of two (32 ... 512) on the walking axis, within a few units in the
    last place of the next integer, so that a step crossing the power of two
    rounds into that integer and (int)_x moves by two cells.  The reference
    walks those float steps; on a 600^2 map of 4000 scattered occupied cells
    every range and every step count equals the oracle's."""
    torch = _torch()
    img = np.zeros((600, 600), np.uint8)
    rng = np.random.default_rng(3)
    for _ in range(4000):
        img[rng.integers(0, 600), rng.integers(0, 600)] = 1
    starts = []
    for p2 in (32.0, 64.0, 128.0, 256.0, 512.0):
        for k in range(1, 6):
            for below in (1.0, 3.0, 7.5):
                starts.append(np.float32(p2 - below - k * p2 * 2.0 ** -24))
    starts = np.array(starts, np.float32)
    # the steps do round into the next integer for some of these starts
    x = starts.copy()
    moved = np.zeros(len(x), bool)
    for _ in range(9):
        nx = (x + np.float32(1.0)).astype(np.float32)
        moved |= np.floor(nx) - np.floor(x) == 2
        x = nx
    assert moved.any()
    across = rng.uniform(5.0, 590.0, 8).astype(np.float32)
    heads = np.array([0.0, 0.3, -0.3, 1.2, 1.5707964, 0.05, -0.05], np.float32)
    q = np.array([p for s in starts for a in across for h in heads for p in ((s, a, h), (a, s, h))], np.float32)
    n = q.shape[0]
    m = rl.DeviceMap(ctx, O.xmajor(img))
    d_q = torch.from_numpy(q).cuda()
    out = torch.full((n,), -1.0, dtype=torch.float32, device="cuda")
    cnt = torch.zeros(n, dtype=torch.int32, device="cuda")
    m.calc_range_many_dev(rl.RL_METHOD_BL, 500.0, d_q.data_ptr(), out.data_ptr(), n)
    m.count_steps_dev(rl.RL_METHOD_BL, rl.RL_FRAME_GRID, 500.0, d_q.data_ptr(), 0, cnt.data_ptr(), n, 1)
    ctx.sync()
    want, want_cnt = O.Oracle(img, 500.0, "bl").calc_range_many(q, counts=True)
    assert_bits_equal(out.cpu().numpy(), want, "ranges")
    assert np.array_equal(cnt.cpu().numpy(), want_cnt)


# --------------------------------------------------------------------------- step counts

def test_step_counts_match_oracle(corridor_map):
    torch = _torch()
    ins = W.world_poses(corridor_map, 500, seed=13)
    ang = W.beam_angles(60)
    m = rl.DeviceMap(rl.default_context(0), O.xmajor(corridor_map))
    d_ins = torch.from_numpy(ins).cuda()
    d_ang = torch.from_numpy(ang).cuda()
    for method, kind in ((rl.RL_METHOD_RM, "rm"), (rl.RL_METHOD_BL, "bl")):
        cnt = torch.zeros(500 * 60, dtype=torch.int32, device="cuda")
        m.count_steps_dev(method, rl.RL_FRAME_ANGLES, 500.0, d_ins.data_ptr(), d_ang.data_ptr(),
                          cnt.data_ptr(), 500, 60)
        m.ctx.sync()
        _, want = O.Oracle(corridor_map, 500.0, kind).numpy_calc_range_angles(ins, ang, counts=True)
        assert np.array_equal(cnt.cpu().numpy(), want)


# --------------------------------------------------------------------------- compact map, fused step

@pytest.mark.parametrize("case", ["rect512", "corridor2048", "open_sparse"])
def test_compact_map_matches_fp32_map_and_oracle(case):
    """The 8-bit index map (with escapes to the fp32 map for distances beyond
    the table) gives the same ranges as the fp32 map and the oracle."""
    img = {"rect512": lambda: W.rect_disk_map(512, seed=0),
           "corridor2048": lambda: W.corridor_map(2048, seed=1),
           # large empty regions: distances up to hundreds of cells -> escapes
           "open_sparse": lambda: (np.random.default_rng(3).random((700, 900)) < 2e-5).astype(np.uint8)}[case]()
    torch = _torch()
    ctx = rl.default_context(0)
    ins = W.world_poses(img, 1500, seed=21)
    ang = W.beam_angles(60)
    want = O.Oracle(img, 500.0, "rm").numpy_calc_range_angles(ins, ang)
    d_ins, d_ang = torch.from_numpy(ins).cuda(), torch.from_numpy(ang).cuda()
    for flags in (0, rl._lib.RL_MAP_NO_COMPACT):
        m = rl.DeviceMap(ctx, O.xmajor(img), flags)
        out = torch.empty(1500 * 60, dtype=torch.float32, device="cuda")
        m.numpy_calc_range_angles_dev(rl.RL_METHOD_RM, 500.0, d_ins.data_ptr(), d_ang.data_ptr(),
                                      out.data_ptr(), 1500, 60)
        ctx.sync()
        assert_bits_equal(out.cpu().numpy(), want, f"{case} flags={flags}")


def test_map_from_precomputed_edt():
    """rl_map_create_from_edt: with the exact transform it equals a map that
    builds its own (and the oracle); with arbitrary values (scaled, negative,
    -0, NaN) the compact march equals the fp32 march over the same values."""
    torch = _torch()
    ctx = rl.default_context(0)
    img = W.rect_disk_map(300, seed=4)
    occ = O.xmajor(img)
    edt = O.edt(img)
    ins = W.world_poses(img, 700, seed=5)
    ang = W.beam_angles(60)
    d_ins, d_ang = torch.from_numpy(ins).cuda(), torch.from_numpy(ang).cuda()

    def cast(dm):
        out = torch.empty(700 * 60, dtype=torch.float32, device="cuda")
        dm.numpy_calc_range_angles_dev(rl.RL_METHOD_RM, 200.0, d_ins.data_ptr(), d_ang.data_ptr(),
                                       out.data_ptr(), 700, 60)
        ctx.sync()
        return out.cpu().numpy()

    want = O.Oracle(img, 200.0, "rm").numpy_calc_range_angles(ins, ang)
    m = rl.DeviceMap(ctx, occ, edt_xmajor=edt)
    assert np.array_equal(m.edt().view(np.uint32), edt.astype(np.float32).view(np.uint32))
    assert_bits_equal(cast(m), want, "precomputed exact transform")
    odd = (edt * np.float32(0.73)).astype(np.float32)
    rng = np.random.default_rng(6)
    pick = rng.random(odd.shape)
    odd[pick < 0.01] = -1.5
    odd[(pick >= 0.01) & (pick < 0.02)] = -0.0
    odd[(pick >= 0.02) & (pick < 0.021)] = np.nan
    a = cast(rl.DeviceMap(ctx, occ, edt_xmajor=odd))
    b = cast(rl.DeviceMap(ctx, occ, rl._lib.RL_MAP_NO_COMPACT, edt_xmajor=odd))
    assert_bits_equal(a, b, "compact vs fp32 march over arbitrary values")
    with pytest.raises(ValueError):
        rl.DeviceMap(ctx, occ, edt_xmajor=edt[:-1])


@pytest.mark.parametrize("kind,m", [("rm", 60), ("rm", 30), ("rm", 100), ("rm", 64), ("rm", 65), ("rm", 128),
                                    ("bl", 60)])
def test_pf_step_bit_exact(kind, m, corridor_map):
    """rl_pf_step_dev == numpy_calc_range_angles then eval_sensor_model."""
    torch = _torch()
    ctx = rl.default_context(0)
    table = W.sensor_table(501)
    n = 3000
    ins = W.world_poses(corridor_map, n, seed=22)
    ang = W.beam_angles(m)
    orc = O.Oracle(corridor_map, 500.0, kind)
    orc.set_sensor_model(table)
    obs = W.noisy_observation(orc.numpy_calc_range_angles(W.world_poses(corridor_map, 1, seed=23), ang), seed=24)
    want_r = orc.numpy_calc_range_angles(ins, ang)
    want_w = orc.eval_sensor_model(obs, want_r, m, n)
    dm = rl.DeviceMap(ctx, O.xmajor(corridor_map))
    dm.set_sensor_table(table)
    d = {k: torch.from_numpy(v).cuda() for k, v in (("ins", ins), ("ang", ang), ("obs", obs))}
    rr = torch.empty(n * m, dtype=torch.float32, device="cuda")
    ww = torch.empty(n, dtype=torch.float64, device="cuda")
    method = rl.RL_METHOD_RM if kind == "rm" else rl.RL_METHOD_BL
    dm.pf_step_dev(method, 500.0, d["ins"].data_ptr(), d["ang"].data_ptr(), d["obs"].data_ptr(),
                   rr.data_ptr(), ww.data_ptr(), n, m)
    ctx.sync()
    assert_bits_equal(rr.cpu().numpy(), want_r, "pf ranges")
    assert np.array_equal(ww.cpu().numpy(), want_w)


@pytest.mark.parametrize("xcd,box", [(0, 0), (0, 1), (1, 0), (1, 1)])
def test_scheduling_options(xcd, box, corridor_map):
    """XCD-aware scheduling and LDS staging change only where data is read from."""
    torch = _torch()
    ctx = rl.Context(0)
    ctx.set_option(rl._lib.RL_OPT_XCD_SCHED, xcd)
    ctx.set_option(rl._lib.RL_OPT_LDS_BOX, box)
    ins = W.world_poses(corridor_map, 2500, seed=31)
    ang = W.beam_angles(45)
    want = O.Oracle(corridor_map, 500.0, "rm").numpy_calc_range_angles(ins, ang)
    m = rl.DeviceMap(ctx, O.xmajor(corridor_map))
    d_ins, d_ang = torch.from_numpy(ins).cuda(), torch.from_numpy(ang).cuda()
    out = torch.empty(2500 * 45, dtype=torch.float32, device="cuda")
    for _ in range(2):  # the region buckets are rebuilt by every launch
        out.fill_(-1)
        m.numpy_calc_range_angles_dev(rl.RL_METHOD_RM, 500.0, d_ins.data_ptr(), d_ang.data_ptr(),
                                      out.data_ptr(), 2500, 45)
        ctx.sync()
        assert_bits_equal(out.cpu().numpy(), want, f"xcd={xcd} box={box}")


def test_sensor_model_nan_observations_and_ranges(corridor_map):
    """NaN observations (no-return beams) and NaN expected ranges index the
    table's max-range row / column on the GPU exactly as in the oracle, in
    eval_sensor_model, the fused call and the particle-filter step."""
    table = W.sensor_table(301)
    n, m, R = 500, 60, 300.0
    ins = W.world_poses(corridor_map, n, seed=31)
    ang = W.beam_angles(m)
    orc = O.Oracle(corridor_map, R, "rm")
    orc.set_sensor_model(table)
    obs = W.noisy_observation(orc.numpy_calc_range_angles(W.world_poses(corridor_map, 1, seed=32), ang), seed=33,
                              max_range=R)
    obs[[0, 7, 59]] = np.nan
    ranges = orc.numpy_calc_range_angles(ins, ang)
    ranges[::97] = np.nan
    meth = rl.PyRayMarchingGPU(rl.PyOMap(corridor_map), R)
    meth.set_sensor_model(table)
    w = np.empty(n, np.float64)
    meth.eval_sensor_model(obs, ranges, w, m, n)
    assert np.array_equal(w, orc.eval_sensor_model(obs, ranges, m, n))
    wf = np.empty(n, np.float64)
    meth.calc_range_repeat_angles_eval_sensor_model(ins, ang, obs, wf)
    assert np.array_equal(wf, orc.calc_range_repeat_angles_eval_sensor_model(ins, ang, obs))


def test_dependent_load_probe(ctx):
    """The latency probe follows the chain it is given: the final index equals
    the host's walk of the same cycle, and the elapsed time is positive."""
    torch = _torch()
    rng = np.random.default_rng(3)
    n = 4096
    perm = rng.permutation(n)
    nxt = np.zeros(n * 32, np.uint32)
    nxt[perm * 32] = (np.roll(perm, -1) * 32).astype(np.uint32)
    d = torch.from_numpy(nxt).cuda()
    out = torch.zeros(2, dtype=torch.int64, device="cuda")
    steps = 1000
    tick_ns = ctx.debug_chase(d.data_ptr(), steps, out.data_ptr())
    torch.cuda.synchronize()
    i = 0
    for _ in range(steps):
        i = int(nxt[i])
    assert int(out[1].item()) == i
    assert int(out[0].item()) > 0 and tick_ns > 0
