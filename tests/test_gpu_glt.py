"""GiantLUTCast on the GPU: the table built on the device against the oracle and
the compiled reference's fixture, and every casting call with RL_METHOD_GLT
bit-exact against the oracle."""
import hashlib
import math
import os

import numpy as np
import pytest

import range_libc_amd as rl
from oracle import oracle as O
from range_libc_amd import workloads as W
from tests.conftest import GOLDEN, assert_bits_equal

pytestmark = pytest.mark.gpu

CASES = {"rect64": (60.0, 108), "corr96x80": (120.0, 37), "rect48": (500.0, 360)}
WORLD = (0.05, 0.3, -3.0, 2.0, math.sin(0.3), math.cos(0.3))


def _torch():
    import torch

    assert torch.cuda.is_available(), "GPU test needs a device"
    return torch


@pytest.fixture(scope="module")
def golden():
    return np.load(os.path.join(GOLDEN, "giant_lut.npz"))


def _map(ctx, img, R, td):
    dm = rl.DeviceMap(ctx, O.xmajor(img))
    dm.build_giant_lut(R, td)
    return dm


@pytest.mark.parametrize("td", [17, 256, 257])
def test_glt_table_ragged_tiles(ctx, td):
    """The whole table of a 70 x 53 map (3710 cells: the table is built in
    tiles of 16 cells, the last one ragged) with 17 and 256 bins (tiled build)
    and 257 (one entry per thread) equals the oracle's."""
    img = np.zeros((53, 70), np.uint8)  # rows = y
    img[[0, -1], :] = 1
    img[:, [0, -1]] = 1
    rng = np.random.default_rng(5)
    img[rng.integers(1, 52, 150), rng.integers(1, 69, 150)] = 1
    dm = _map(ctx, img, 40.0, td)
    table = dm.giant_lut()[0]
    assert np.array_equal(table, O.Oracle(img, 40.0, "glt", theta_disc=td).glt)


@pytest.mark.parametrize("name", sorted(CASES))
def test_glt_table_and_queries_golden(ctx, golden, name):
    R, td = CASES[name]
    img = golden[f"{name}_img"]
    dm = _map(ctx, img, R, td)
    table, td_got, r_got, ms = dm.giant_lut()
    assert td_got == td and r_got == np.float32(R) and ms > 0
    assert hashlib.sha256(np.ascontiguousarray(table).tobytes()).hexdigest() == str(golden[f"{name}_table_sha"])
    assert np.array_equal(table, O.Oracle(img, R, "glt", theta_disc=td).glt)
    q = golden[f"{name}_ins"]
    n = q.shape[0]
    out = np.empty(n, np.float32)
    dm.calc_range_many(rl.RL_METHOD_GLT, R, q, out, n)
    assert_bits_equal(out, golden[f"{name}_grid"], "grid")
    dm.numpy_calc_range(rl.RL_METHOD_GLT, R, q, out, n)
    assert_bits_equal(out, golden[f"{name}_world"], "world")
    ang = W.beam_angles(61)
    outa = np.empty(300 * 61, np.float32)
    dm.numpy_calc_range_angles(rl.RL_METHOD_GLT, R, q[:300], ang, outa, 300, 61)
    assert_bits_equal(outa, golden[f"{name}_angles"], "angles")
    dm.set_world(*WORLD)
    dm.numpy_calc_range(rl.RL_METHOD_GLT, R, q, out, n)
    assert_bits_equal(out, golden[f"{name}_world_rot"], "rotated world")


def test_glt_sensor_fused_and_pf_step(ctx):
    torch = _torch()
    img = W.corridor_map(200, seed=6)
    R, td = 150.0, 90
    dm = _map(ctx, img, R, td)
    table = W.sensor_table(151)
    dm.set_sensor_table(table)
    o = O.Oracle(img, R, "glt", theta_disc=td)
    o.set_sensor_model(table)
    ins = W.world_poses(img, 333, seed=2)
    ang = W.beam_angles(60)
    want_r = o.numpy_calc_range_angles(ins, ang)
    obs = W.noisy_observation(want_r[:60], seed=1, max_range=R)
    want_w = o.eval_sensor_model(obs, want_r, 60, 333)
    want_fused = o.calc_range_repeat_angles_eval_sensor_model(ins, ang, obs)
    w = np.empty(333, np.float64)
    dm.repeat_angles_eval_sensor_model(rl.RL_METHOD_GLT, R, ins, ang, obs, w, 333, 60)
    assert np.array_equal(w, want_fused)
    d = {k: torch.from_numpy(v).cuda() for k, v in (("ins", ins), ("ang", ang), ("obs", obs))}
    rng_ = torch.empty(333 * 60, dtype=torch.float32, device="cuda")
    wd = torch.empty(333, dtype=torch.float64, device="cuda")
    dm.pf_step_dev(rl.RL_METHOD_GLT, R, d["ins"].data_ptr(), d["ang"].data_ptr(), d["obs"].data_ptr(),
                   rng_.data_ptr(), wd.data_ptr(), 333, 60)
    ctx.sync()
    assert_bits_equal(rng_.cpu().numpy(), want_r, "pf_step ranges")
    assert np.array_equal(wd.cpu().numpy(), want_w)


def test_glt_batch(ctx):
    torch = _torch()
    jobs, keep, want = [], [], []
    for i, (side, R, td) in enumerate([(70, 60.0, 40), (120, 200.0, 72), (90, 90.0, 16)]):
        img = W.rect_disk_map(side, seed=i)
        dm = _map(ctx, img, R, td)
        ins = W.world_poses(img, 50 + 30 * i, seed=i)
        ang = W.beam_angles(17 + 20 * i)
        want.append(O.Oracle(img, R, "glt", theta_disc=td).numpy_calc_range_angles(ins, ang))
        d_ins, d_ang = torch.from_numpy(ins).cuda(), torch.from_numpy(ang).cuda()
        out = torch.full((ins.shape[0] * ang.shape[0],), -1.0, dtype=torch.float32, device="cuda")
        keep.append((dm, d_ins, d_ang, out))
        jobs.append({"map": dm, "max_range": R, "ins": d_ins.data_ptr(), "angles": d_ang.data_ptr(),
                     "ranges": out.data_ptr(), "num_particles": ins.shape[0], "num_angles": ang.shape[0]})
    b = rl.Batch(ctx, rl.RL_METHOD_GLT)
    b.set_jobs(jobs)
    b.launch()
    ctx.sync()
    for k, (w, (_, _, _, out)) in enumerate(zip(want, keep)):
        assert_bits_equal(out.cpu().numpy(), w, f"job {k}")


def test_py_giant_lut_cast(ctx):
    img = W.rect_disk_map(80, seed=2)
    omap = rl.PyOMap(img)
    g = rl.PyGiantLUTCast(omap, 100.0, 64)
    assert g.lut_size() == 80 * 80 * 64 * 2
    o = O.Oracle(img, 100.0, "glt", theta_disc=64)
    assert g.calc_range(10.5, 20.25, 1.0) == o.calc_range(10.5, 20.25, 1.0)
    ins = W.world_poses(img, 100, seed=3)
    out = np.empty(100, np.float32)
    g.calc_range_many(ins, out)
    assert_bits_equal(out, o.numpy_calc_range(ins), "calc_range_many")
    ang = W.beam_angles(10)
    outa = np.empty(1000, np.float32)
    g.calc_range_repeat_angles(ins, ang, outa)
    assert_bits_equal(outa, o.numpy_calc_range_angles(ins, ang), "repeat angles")


def test_glt_special_headings_and_errors(ctx):
    img = W.rect_disk_map(50, seed=8)
    R, td = 80.0, 50
    dm = rl.DeviceMap(ctx, O.xmajor(img))
    q = np.array([[10.5, 10.5, 1.0]], np.float32)
    out = np.empty(1, np.float32)
    with pytest.raises(rl.RangeLibError):  # no table yet
        dm.calc_range_many(rl.RL_METHOD_GLT, R, q, out, 1)
    dm.build_giant_lut(R, td)
    with pytest.raises(rl.RangeLibError):  # table built for another max_range
        dm.calc_range_many(rl.RL_METHOD_GLT, R + 1.0, q, out, 1)
    with pytest.raises(rl.RangeLibError):
        dm.build_giant_lut(R, 0)
    o = O.Oracle(img, R, "glt", theta_disc=td)
    heads = np.array([float("nan"), 1e10, -1e10, 3e4, -3e4, 6.28318530718, -1e-30, 0.0, -0.0], np.float32)
    q = np.stack([np.full(heads.shape, 12.5, np.float32), np.full(heads.shape, 20.5, np.float32), heads], 1)
    q = np.concatenate([q, np.array([[float("nan"), 3.0, 1.0], [3.0, float("inf"), 1.0], [-0.5, 3.0, 1.0]],
                                    np.float32)])
    out = np.empty(q.shape[0], np.float32)
    dm.calc_range_many(rl.RL_METHOD_GLT, R, q, out, q.shape[0])
    want = o.calc_range_many(q)
    assert np.isnan(out[0]) and np.isnan(out[1]) and np.isnan(out[2])  # NaN / non-terminating headings
    assert_bits_equal(out[3:9], want[3:9], "finite headings")
    assert np.isnan(out[9]) and out[10] == np.float32(R) and out[11] == np.float32(R)
