"""Multi-map batches: every launch configuration bit-exact against the oracle,
the autotuner, and HIP-graph capture of the batch launch."""
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
            for order in (0, 1):
                for grid in (0, 5):
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
