"""Multi-map batches: every launch configuration bit-exact against the oracle,
the autotuner, and HIP-graph capture of the batch launch."""
import os

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


def _make_jobs(ctx, specs, seed):
    """Per map: device buffers and the oracle's expected ranges / weights."""
    torch = _torch()
    rng = np.random.default_rng(seed)
    table = W.sensor_table(301)
    jobs, want, keep = [], [], []
    for i, sp in enumerate(specs):
        img = W.make_map(sp["kind"], sp["side"], sp["map_seed"])
        n, m, R = sp["particles"], sp["beams"], float(sp["max_range"])
        ins = W.world_poses(img, n, seed=sp["pose_seed"])
        ang = W.beam_angles(m)
        orc = O.Oracle(img, R, "rm")
        orc.set_sensor_model(table)
        obs = W.noisy_observation(orc.numpy_calc_range_angles(W.world_poses(img, 1, seed=i + 100), ang),
                                  seed=i, max_range=R)
        wr = orc.numpy_calc_range_angles(ins, ang)
        sensor = bool(sp.get("sensor", True))
        ww = orc.eval_sensor_model(obs, wr, m, n) if sensor else None
        dm = rl.DeviceMap(ctx, O.xmajor(img))
        dm.set_sensor_table(table)
        d = {k: torch.from_numpy(v).cuda() for k, v in (("ins", ins), ("ang", ang), ("obs", obs))}
        d["rng"] = torch.full((n * m,), -1.0, dtype=torch.float32, device="cuda")
        d["w"] = torch.full((n,), -1.0, dtype=torch.float64, device="cuda")
        jobs.append({"map": dm, "max_range": R, "ins": d["ins"].data_ptr(), "angles": d["ang"].data_ptr(),
                     "obs": d["obs"].data_ptr() if sensor else 0, "ranges": d["rng"].data_ptr(),
                     "weights": d["w"].data_ptr() if sensor else 0, "num_particles": n, "num_angles": m})
        want.append((wr, ww))
        keep.append((dm, d))
    return jobs, want, keep


def _specs():
    rng = np.random.default_rng(5)
    specs = []
    for i in range(7):
        specs.append({"kind": "corridor" if i % 2 else "rect_disk", "side": int(rng.integers(64, 700)),
                      "map_seed": 10 + i, "particles": int(rng.integers(1, 300)),
                      "beams": int([1, 7, 30, 60, 64, 65, 120][i]), "max_range": [500, 100, 300, 500, 50, 500, 200][i],
                      "pose_seed": 20 + i, "sensor": i != 2})
    return specs


def _check(keep, want):
    for k, ((dm, d), (wr, ww)) in enumerate(zip(keep, want)):
        assert_bits_equal(d["rng"].cpu().numpy(), wr, f"job {k} ranges")
        if ww is not None:
            assert np.array_equal(d["w"].cpu().numpy(), ww), f"job {k} weights"


def test_batch_every_config_bit_exact():
    ctx = rl.Context(0)
    jobs, want, keep = _make_jobs(ctx, _specs(), 1)
    b = rl.Batch(ctx)
    b.set_jobs(jobs)
    for rpw in (64, 32, 16):
        for box in ((0, 1) if rpw == 64 else (0,)):
            for order in (0, 1, 2):
                for grid in (0, 5, 16):
                    for _, d in keep:
                        d["rng"].fill_(-1)
                        d["w"].fill_(-1)
                    b.set_config(rpw, box, order, grid)
                    b.launch()
                    ctx.sync()
                    _check(keep, want)


def test_batch_autotune_and_graph_capture():
    torch = _torch()
    ctx = rl.Context(0)
    stream = torch.cuda.Stream()
    ctx.set_stream(stream.cuda_stream)
    with torch.cuda.stream(stream):
        jobs, want, keep = _make_jobs(ctx, _specs(), 2)
        b = rl.Batch(ctx)
        b.set_jobs(jobs)
        ms = b.autotune(reps=1)
        assert ms > 0 and b.config()["rays_per_wave"] in (16, 32, 64)
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g, stream=stream):
            b.launch()
        for rep in range(3):
            for _, d in keep:  # eager work between replays
                d["rng"].fill_(-1)
                d["w"].fill_(-1)
            g.replay()
            stream.synchronize()
            _check(keep, want)
    ctx.set_stream(None)


def test_batch_rejects_bad_jobs():
    ctx = rl.Context(0)
    b = rl.Batch(ctx)
    dm = rl.DeviceMap(ctx, np.zeros((8, 8), np.uint8))
    with pytest.raises(rl.RangeLibError):
        b.set_jobs([{"map": dm, "max_range": 10.0, "ins": 0, "angles": 0, "ranges": 0,
                     "num_particles": 1, "num_angles": 1}])


def test_batch_follows_map_changes_and_keeps_maps_alive():
    """A setter that replaces a job map's sensor table or world frame between
    launches is seen by the next launch (the descriptors are uploaded again),
    and a map whose caller handle is closed stays valid while a batch holds it."""
    ctx = rl.Context(0)
    specs = _specs()[:3]
    jobs, want, keep = _make_jobs(ctx, specs, 3)
    b = rl.Batch(ctx)
    b.set_jobs(jobs)
    b.launch()
    ctx.sync()
    _check(keep, want)
    # job 0: another sensor table (a new device buffer) and a rotated, scaled world frame
    sp = specs[0]
    img = W.make_map(sp["kind"], sp["side"], sp["map_seed"])
    n, m, R = sp["particles"], sp["beams"], float(sp["max_range"])
    table2 = W.sensor_table(401, sigma_hit=5.0)
    world = (0.5, 0.3, 1.5, -2.0)
    import math
    world6 = (world[0], world[1], world[2], world[3], math.sin(world[1]), math.cos(world[1]))
    dm0, d0 = keep[0]
    dm0.set_sensor_table(table2)
    dm0.set_world(*world6)
    orc = O.Oracle(img, R, "rm", world=world6)
    orc.set_sensor_model(table2)
    ins = d0["ins"].cpu().numpy()
    ang = d0["ang"].cpu().numpy()
    obs = d0["obs"].cpu().numpy()
    wr = orc.numpy_calc_range_angles(ins, ang)
    want[0] = (wr, orc.eval_sensor_model(obs, wr, m, n))
    # job 1's map: the caller's handle goes away; the batch keeps the map
    keep[1][0].close()
    for _, d in keep:
        d["rng"].fill_(-1)
        d["w"].fill_(-1)
    b.launch()
    ctx.sync()
    _check(keep, want)
    b.close()


def _c4_job(sp, table, i, R=500.0):
    """One map of the 64-map workload with the oracle's ranges and weights."""
    img = W.make_map(sp["kind"], sp["side"], sp["map_seed"])
    n, m = sp["particles"], sp["beams"]
    ins = W.world_poses(img, n, seed=sp["pose_seed"])
    ang = W.beam_angles(m)
    orc = O.Oracle(img, R, "rm")
    orc.set_sensor_model(table)
    obs = W.noisy_observation(orc.numpy_calc_range_angles(W.world_poses(img, 1, seed=sp["pose_seed"] + 1), ang),
                              seed=i, max_range=R)
    wr = orc.numpy_calc_range_angles(ins, ang)
    ww = orc.eval_sensor_model(obs, wr, m, n)
    del orc
    return img, ins, ang, obs, wr, ww


def test_batch_uneven64_full_size():
    """The 64-map workload at its full size: maps of side 256-4096 (both
    generators; sides above 2896 take the distance transform's exact-search
    fallback rows), 500-8000 particles, 30-120 beams, max_range 500, 20.75 M
    rays with the sensor model.  One launch with the autotuned configuration
    and one with a fixed one; every range and every weight is compared bit for
    bit with the oracle."""
    from concurrent.futures import ThreadPoolExecutor

    torch = _torch()
    specs = W.uneven_batch_specs(64, seed=4)
    assert max(sp["side"] for sp in specs) > 2896
    table = W.sensor_table(501)
    print(f"c4: building 64 maps and the oracle's outputs", flush=True)
    with ThreadPoolExecutor(max_workers=min(16, os.cpu_count() or 1)) as ex:
        data = list(ex.map(lambda a: _c4_job(a[1], table, a[0]), enumerate(specs)))
    ctx = rl.Context(0)
    jobs, keep = [], []
    total = 0
    for sp, (img, ins, ang, obs, wr, ww) in zip(specs, data):
        n, m = sp["particles"], sp["beams"]
        dm = rl.DeviceMap(ctx, O.xmajor(img))
        dm.set_sensor_table(table)
        d = {k: torch.from_numpy(v).cuda() for k, v in (("ins", ins), ("ang", ang), ("obs", obs))}
        d["rng"] = torch.full((n * m,), -1.0, dtype=torch.float32, device="cuda")
        d["w"] = torch.full((n,), -1.0, dtype=torch.float64, device="cuda")
        jobs.append({"map": dm, "max_range": 500.0, "ins": d["ins"].data_ptr(), "angles": d["ang"].data_ptr(),
                     "obs": d["obs"].data_ptr(), "ranges": d["rng"].data_ptr(), "weights": d["w"].data_ptr(),
                     "num_particles": n, "num_angles": m})
        keep.append((dm, d))
        total += n * m
    assert total > 20_000_000
    want = [(wr, ww) for (_, _, _, _, wr, ww) in data]
    b = rl.Batch(ctx)
    b.set_jobs(jobs)
    b.autotune(reps=1)
    print(f"c4: autotuned {b.config()}", flush=True)
    for cfg in (None, (64, 0, 1, 0)):
        if cfg is not None:
            b.set_config(*cfg)
        for _, d in keep:
            d["rng"].fill_(-1)
            d["w"].fill_(-1)
        b.launch()
        ctx.sync()
        _check(keep, want)
    b.close()
