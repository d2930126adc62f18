"""Benchmark: one particle-filter step on one MI355X.

Workload (default): 2048x2048 synthetic corridor map, 4000 particles in free
space, ``calc_range_repeat_angles`` with 60 beams over 270 degrees and
max_range 500 px (240,000 rays), then ``eval_sensor_model`` with a 501x501
table against 60 observed ranges.  Inputs are device-resident when the timed
region starts; a step is one fused launch (ray marching over repeated angles
with the sensor-model product of each particle), bit-identical to the two
calls, which are timed beside it.

Prints one JSON line: rays/s, the fused kernel's fraction of the HBM roofline
(algorithmic bytes = 4 B per distance-transform read, counted per ray on the
device, plus pose/angle/observation input, range output, one sensor-table entry
per ray and one weight per particle), the longest dependent chain of reads of
the launch, and the reference CPU path timed on this host.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
    python bench.py --workload {c1,uneven64,surrogate,glt,cddt,bresenham}
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, published
MAP_SIDE = 2048
PARTICLES = 4000
BEAMS = 60
MAX_RANGE = 500.0
TABLE = 501


def build_workload(seed_map=1, seed_pose=2, seed_true=3, seed_noise=4):
    from range_libc_amd import workloads as W

    img = W.corridor_map(MAP_SIDE, seed_map)
    ins = W.world_poses(img, PARTICLES, seed_pose)
    angles = W.beam_angles(BEAMS)
    table = W.sensor_table(TABLE)
    true_pose = W.world_poses(img, 1, seed_true)
    return img, ins, angles, table, true_pose, seed_noise


def host_info() -> dict:
    """CPU model, usable cores and C library of the host the CPU leg runs on."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        libc = os.confstr("CS_GNU_LIBC_VERSION")
    except (ValueError, OSError):
        libc = "unknown"
    return {"cpu_model": model, "host_cores": len(os.sched_getaffinity(0)), "libc": libc}


def timed_runs(fn, budget_s: float, min_runs: int = 10, warmup: int = 2):
    """Wall times of repeated runs of fn: `warmup` untimed runs, then at least
    `min_runs` timed ones and more while the budget lasts."""
    for _ in range(warmup):
        fn()
    times = []
    t_end = time.perf_counter() + budget_s
    while len(times) < min_runs or time.perf_counter() < t_end:
        t0 = time.perf_counter()
        fn()
        times.append(time.perf_counter() - t0)
    return times


def rate_entry(units_per_run: float, times, unit: str, cores: int, kind: str, sample: str) -> dict:
    med = float(np.median(times))
    return {"value": units_per_run / med, "unit": unit, "cores": cores, "kind": kind,
            "sample": f"{sample}; median of {len(times)} runs ({med * 1e3:.3f} ms each) after 2 warm-ups",
            **host_info()}


def cpu_baseline(img, ins, angles, obs, budget_s=10.0):
    """The reference's CPU path (oracle/_ref, or the oracle port when the
    compiled reference is absent) on this host's cores: whole particle-filter
    steps (numpy_calc_range_angles + eval_sensor_model) split over threads,
    repeated until ~budget_s of wall time."""
    from oracle import oracle as O
    from range_libc_amd import workloads as W

    threads = len(os.sched_getaffinity(0))
    table = W.sensor_table(TABLE)
    if O.ref_available():
        kind = "reference"
        m = O.Reference(img, MAX_RANGE, "rm")
        m.set_sensor_model(table)
        step = lambda: m.pf_step(ins, angles, obs, threads)  # noqa: E731
    else:
        kind = "port"
        threads = 1
        m = O.Oracle(img, MAX_RANGE, "rm")
        m.set_sensor_model(table)

        def step():
            r = m.numpy_calc_range_angles(ins, angles)
            return m.eval_sensor_model(obs, r, BEAMS, PARTICLES)
    rays = PARTICLES * BEAMS
    times = timed_runs(step, budget_s)
    out = rate_entry(rays, times, "rays/s", threads, kind,
                     f"full steps ({PARTICLES}x{BEAMS} rays + sensor model) on {threads} threads")
    if kind == "reference" and threads > 1:  # the reference's own single-threaded path as well
        t1 = timed_runs(lambda: m.pf_step(ins, angles, obs, 1), budget_s / 2)
        out["single_thread"] = rate_entry(rays, t1, "rays/s", 1, kind, "full steps on 1 thread")
    return out


def _stream_ctx():
    import torch

    import range_libc_amd as rl

    dev = torch.device("cuda", 0)
    stream = torch.cuda.Stream(device=dev)
    torch.cuda.set_stream(stream)
    ctx = rl.default_context(0)
    ctx.set_stream(stream.cuda_stream)
    return torch, rl, dev, stream, ctx


def _kernel_ms(torch, stream, fn, reps):
    """Mean time per launch of `reps` back-to-back launches of fn on `stream`,
    bracketed by one HIP-event pair behind a spin kernel (no host gap enters)."""
    fn()
    torch.cuda._sleep(50_000_000)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(reps):
        fn()
    e1.record(stream)
    stream.synchronize()
    return e0.elapsed_time(e1) / reps


def dependent_load_ns(torch, ctx, dev, steps=20000):
    """ns per dependent load of one lane chasing a random cycle, one element per
    128-byte line: through 2 MiB after a full warm pass (held in one XCD's
    4 MiB L2), through the first lines of a 64 MiB cycle after a full warm pass
    (evicted from L2 by the rest of the pass, held in the 256 MiB Infinity
    Cache), and through 2 GiB never read before (HBM).  Timed with HIP events
    around one launch of `steps` loads on the library's stream (launch
    overhead below 0.1%)."""
    out = {}
    rng = np.random.default_rng(0)
    res = torch.zeros(2, dtype=torch.int64, device=dev)
    stream = torch.cuda.current_stream()
    for name, size, warm in (("l2_2MiB", 2 << 20, True), ("mall_64MiB", 64 << 20, True), ("hbm_2GiB", 2 << 30, False)):
        n = size // 128
        perm = rng.permutation(n).astype(np.int64)
        nxt = np.zeros(size // 4, np.uint32)
        nxt[perm * 32] = (np.roll(perm, -1) * 32).astype(np.uint32)
        del perm
        d = torch.from_numpy(nxt).to(dev)
        del nxt
        ctx.debug_chase(d.data_ptr(), 1, res.data_ptr())  # code and first line loaded
        if warm:
            ctx.debug_chase(d.data_ptr(), max(n, steps), res.data_ptr())
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        ctx.debug_chase(d.data_ptr(), steps, res.data_ptr())
        e1.record(stream)
        stream.synchronize()
        out[name] = e0.elapsed_time(e1) * 1e6 / steps
        del d
        torch.cuda.empty_cache()
    return out


def _march_seq(edt, x0, y0, dx, dy, max_range):
    """Distances read by the march of one ray, in order (float32 as the kernel)."""
    f32, one, c999 = np.float32, np.float32(1.0), np.float32(0.999)
    x0, y0, dx, dy = f32(x0), f32(y0), f32(dx), f32(dy)
    seq = []
    t = f32(0.0)
    while t < max_range:
        px, py = int(x0 + dx * t), int(y0 + dy * t)
        if not (0 <= px < edt.shape[0] and 0 <= py < edt.shape[1]):
            break
        d = edt[px, py]
        seq.append(d)
        if d <= 0:
            break
        st = d * c999
        t = f32(t + (one if st < one else st))
    return seq


_TABULATED = None


def _rounds(seq, start_k=0, prev=None, width=2, limit=None):
    """Rounds of the march schedule over a read sequence from position start_k:
    one read per round for the first 8, then speculation of `width` positions
    (a round also resolves the positions after cells that repeat the previous
    tabulated distance).  Stops after `limit` rounds; returns (rounds, k, prev)."""
    global _TABULATED
    if _TABULATED is None:
        sq = sorted({a * a + b * b for a in range(32) for b in range(a, 32) if a * a + b * b < 1024})[:255]
        _TABULATED = {np.float32(np.sqrt(np.float32(v))) for v in sq}
    rounds, k = 0, start_k
    while k < len(seq) and (limit is None or rounds < limit):
        if k < 8 or prev is None or prev not in _TABULATED:
            k += 1
        else:
            j = 0
            while j < width - 1 and k + j < len(seq) and seq[k + j] == prev:
                j += 1
            k = min(k + j + 1, len(seq))
        prev = seq[k - 1]
        rounds += 1
    return rounds, k, prev


def longest_chain_rounds(m, ins, angles, counts, ctx, torch, dev, top=16):
    """Dependent read rounds of the longest waves under the particle kernel's
    schedule, replayed from the device's own distance map and sinf/cosf: every
    ray of the wave takes width-2 rounds; once one ray is left, the wave takes
    it on with 8-wide speculation.  Returns the most rounds among the `top`
    waves whose rays read the most cells."""
    edt = m.edt()
    M = angles.size
    parts = np.argsort(counts.reshape(-1, M).max(1))[::-1][:top]
    rot = np.float32(-0.0 - 3 * np.pi / 2)
    heads = ((-ins[parts, 2] + rot).astype(np.float32)[:, None] - angles[None, :]).astype(np.float32).ravel()
    dh = torch.from_numpy(heads).to(dev)
    ds, dc = torch.empty_like(dh), torch.empty_like(dh)
    ctx.sincosf_dev(dh.data_ptr(), ds.data_ptr(), dc.data_ptr(), heads.size)
    torch.cuda.synchronize()
    sn, cs = ds.cpu().numpy().reshape(-1, M), dc.cpu().numpy().reshape(-1, M)
    best = 0
    for i, p in enumerate(parts):
        seqs = [_march_seq(edt, ins[p, 1], ins[p, 0], cs[i, a], sn[i, a], MAX_RANGE) for a in range(M)]
        r = sorted((_rounds(q)[0], a) for a, q in enumerate(seqs))
        r2, (_, a1) = (r[-2][0] if len(r) > 1 else 0), r[-1]
        n1, k, prev = _rounds(seqs[a1], limit=r2)
        n2, _, _ = _rounds(seqs[a1], start_k=k, prev=prev, width=8)
        best = max(best, n1 + n2)
    return best


def bench_c1(args):
    """The reference's CPU-runnable configuration: plain ray marching through
    numpy_calc_range (RangeLib.h:439-480, 927-962) on a seeded 512x512 map of 25
    rectangles and 15 disks with a 2-px border, 100,000 world-frame queries from
    uniformly drawn free cells (sub-cell offsets, headings uniform in [0, 2 pi)),
    max_range 500.  One step = one launch over the 100,000 device-resident
    queries; every range is compared with the oracle."""
    from oracle import oracle as O
    from range_libc_amd import workloads as W

    torch, rl, dev, stream, ctx = _stream_ctx()
    N, R = 100_000, 500.0
    img = W.rect_disk_map(512, seed=0)
    q = W.world_poses(img, N, seed=5)
    occ = O.xmajor(img)
    rl.DeviceMap(ctx, occ[:64, :64].copy())  # EDT kernels' code loaded before the timed build
    m = rl.DeviceMap(ctx, occ)
    edt_ms = m.edt_build_ms()
    d_q = torch.from_numpy(q).to(dev)
    d_out = torch.empty(N, dtype=torch.float32, device=dev)
    d_cnt = torch.zeros(N, dtype=torch.int32, device=dev)
    m.count_steps_dev(rl.RL_METHOD_RM, rl.RL_FRAME_WORLD, R, d_q.data_ptr(), 0, d_cnt.data_ptr(), N)
    stream.synchronize()
    reads = int(d_cnt.sum().item())
    max_reads = int(d_cnt.max().item())
    alg = 4 * reads + 16 * N  # 4 B per distance-map read, 12 B pose in, 4 B range out

    def cast():
        m.numpy_calc_range_dev(rl.RL_METHOD_RM, R, d_q.data_ptr(), d_out.data_ptr(), N)

    for _ in range(args.warmup):
        cast()
    kern_ms = _kernel_ms(torch, stream, cast, args.steps)
    got = d_out.cpu().numpy()
    want = O.Oracle(img, R, "rm").numpy_calc_range(q)
    mismatches = int(np.sum(got.view(np.uint32) != want.view(np.uint32)))
    h_out = np.empty(N, np.float32)
    host_t = timed_runs(lambda: m.numpy_calc_range(rl.RL_METHOD_RM, R, q, h_out, N), 1.0, min_runs=10)
    assert np.array_equal(h_out.view(np.uint32), got.view(np.uint32))
    achieved = alg / (kern_ms * 1e-3) / 1e9
    result = {
        "metric": "rays/s on 1 GPU (numpy_calc_range, plain ray marching, 512^2 rectangles+disks, 100k queries)",
        "value": N / (kern_ms * 1e-3), "unit": "rays/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": kern_ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "c1_rect_disk512_100k_r500", "map": "512x512 rectangles+disks (seed 0)",
                   "queries": N, "max_range": R, "parallelism": "single_gpu"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None, "kernel": "k_cast<RM, world>",
                     "kernel_ms": kern_ms, "alg_bytes_per_launch": alg, "edt_reads_per_ray": reads / N,
                     "max_reads_per_ray": max_reads},
        "mismatches_vs_oracle": mismatches, "rays_checked": N,
        "host_api_ms_per_call": float(np.median(host_t)) * 1e3,
        "host_api_rays_per_s": N / float(np.median(host_t)),
        "edt_build_ms": edt_ms,
    }
    if not args.no_cpu and O.ref_available():
        threads = len(os.sched_getaffinity(0))
        ref = O.Reference(img, R, "rm")
        tm = timed_runs(lambda: ref.numpy_calc_range(q, threads), args.cpu_budget / 2)
        t1 = timed_runs(lambda: ref.numpy_calc_range(q, 1), args.cpu_budget / 2)
        result["cpu_baseline"] = rate_entry(N, tm, "rays/s", threads, "reference",
                                            f"numpy_calc_range over all {N} queries on {threads} threads")
        result["cpu_baseline"]["single_thread"] = rate_entry(N, t1, "rays/s", 1, "reference",
                                                             f"numpy_calc_range over all {N} queries on 1 thread")
        te = timed_runs(lambda: O.ref_edt(img), 2.0, min_runs=5, warmup=1)
        result["cpu_baseline"]["edt_build_ms"] = float(np.median(te)) * 1e3
    return result


def bench_uneven64(args):
    """Secondary workload: 64 maps of uneven size (side log-uniform in
    [256, 4096], rectangle/disk or corridor generator), 500-8000 particles and
    30-120 beams over 270 degrees each, max_range 500, per-particle sensor
    model -- all cast in one autotuned launch, compared with one fused launch
    per map."""
    import torch

    import range_libc_amd as rl
    from range_libc_amd import workloads as W

    dev = torch.device("cuda", 0)
    stream = torch.cuda.Stream(device=dev)
    torch.cuda.set_stream(stream)
    ctx = rl.default_context(0)
    ctx.set_stream(stream.cuda_stream)
    table = W.sensor_table(TABLE)
    specs = W.uneven_batch_specs(64, seed=4)
    t0 = time.perf_counter()
    jobs, keep = [], []
    total_rays = 0
    for i, sp in enumerate(specs):
        img = W.make_map(sp["kind"], sp["side"], sp["map_seed"])
        n, m = sp["particles"], sp["beams"]
        dm = rl.DeviceMap(ctx, np.ascontiguousarray(img.T))
        dm.set_sensor_table(table)
        ins = torch.from_numpy(W.world_poses(img, n, sp["pose_seed"])).to(dev)
        ang = torch.from_numpy(W.beam_angles(m)).to(dev)
        true = torch.from_numpy(W.world_poses(img, 1, sp["pose_seed"] + 1)).to(dev)
        tr = torch.empty(m, device=dev)
        dm.numpy_calc_range_angles_dev(rl.RL_METHOD_RM, MAX_RANGE, true.data_ptr(), ang.data_ptr(), tr.data_ptr(), 1, m)
        stream.synchronize()
        obs = torch.from_numpy(W.noisy_observation(tr.cpu().numpy(), i, max_range=MAX_RANGE)).to(dev)
        rng_ = torch.empty(n * m, device=dev)
        w = torch.empty(n, dtype=torch.float64, device=dev)
        jobs.append({"map": dm, "max_range": MAX_RANGE, "ins": ins.data_ptr(), "angles": ang.data_ptr(),
                     "obs": obs.data_ptr(), "ranges": rng_.data_ptr(), "weights": w.data_ptr(),
                     "num_particles": n, "num_angles": m})
        keep.append((dm, ins, ang, obs, rng_, w))
        total_rays += n * m
    setup_s = time.perf_counter() - t0
    batch = rl.Batch(ctx)
    batch.set_jobs(jobs)
    if args.batch_config:  # a fixed configuration (counter passes over one configuration)
        rpw, box, order, grid = (int(v) for v in args.batch_config.split(","))
        batch.set_config(rays_per_wave=rpw, lds_box=box, order=order, grid=grid)
        tuned_ms = None
    else:
        tuned_ms = batch.autotune(reps=3)

    def timed(fn, reps):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda._sleep(50_000_000)
        e0.record(stream)
        for _ in range(reps):
            fn()
        e1.record(stream)
        stream.synchronize()
        return e0.elapsed_time(e1) / reps

    batch.launch()
    batch_ms = timed(batch.launch, args.steps)

    def per_map():
        for j in jobs:
            j["map"].pf_step_dev(rl.RL_METHOD_RM, MAX_RANGE, j["ins"], j["angles"], j["obs"], j["ranges"],
                                 j["weights"], j["num_particles"], j["num_angles"])

    per_map()
    per_map_ms = timed(per_map, max(1, args.steps // 4))
    # algorithmic bytes of one batch launch: per job 4 B per distance-map read
    # (counted on the device), 12 B per pose, 8 B per beam (angle, observation),
    # 4 B range out and 8 B table entry per ray, 8 B weight per particle
    reads = 0
    max_reads = 0
    for j, (dm, ins, ang, obs, rng_, w) in zip(jobs, keep):
        n, mm = j["num_particles"], j["num_angles"]
        cnt = torch.zeros(n * mm, dtype=torch.int32, device=dev)
        # num_particles is the pose count: the call writes n * mm counts
        dm.count_steps_dev(rl.RL_METHOD_RM, rl.RL_FRAME_ANGLES, MAX_RANGE, ins.data_ptr(), ang.data_ptr(),
                           cnt.data_ptr(), n, mm)
        stream.synchronize()
        reads += int(cnt.sum().item())
        max_reads = max(max_reads, int(cnt.max().item()))
    alg = 4 * reads + sum(12 * j["num_particles"] + 8 * j["num_angles"] + 12 * j["num_particles"] * j["num_angles"]
                          + 8 * j["num_particles"] for j in jobs)
    achieved = alg / (batch_ms * 1e-3) / 1e9
    result = {
        "metric": "rays/s on 1 GPU (uneven 64-map batch, one launch)", "value": total_rays / (batch_ms * 1e-3),
        "unit": "rays/s", "ms_per_step": batch_ms, "per_map_launches_ms": per_map_ms,
        "speedup_vs_per_map": per_map_ms / batch_ms, "autotuned": batch.config(), "autotune_best_ms": tuned_ms,
        "rays": total_rays, "maps": len(jobs), "setup_s": setup_s, "steps": args.steps,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None, "kernel": "k_batch<RM>",
                     "kernel_ms": batch_ms, "alg_bytes_per_launch": alg, "edt_reads_per_ray": reads / total_rays,
                     "max_reads_per_ray": max_reads},
        "config": {"workload": "uneven64_side256-4096_p500-8000_b30-120_r500"}}
    # HBM traffic and instruction issue of the launch, from the committed counter passes
    tpath = os.path.join(ROOT, "profiles", "traffic_batch_uneven64.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            tj = json.load(f)
        result["roofline"]["traffic"] = tj["hbm_bytes_per_launch"]
        result["roofline"]["traffic_source"] = (
            f"profiles/traffic_batch_uneven64.json: rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE passes over this "
            f"launch in configuration {tj['config']} (profiled launches {tj['profiled_launch_us']} us), scaled by "
            f"a 256 MiB copy's counts; not measured in this run")
        result["issue_bound"] = {"valu_us": tj["issue"]["valu_issue_bound_us"],
                                 "salu_us": tj["issue"]["salu_issue_bound_us"],
                                 "wait_frac": tj["issue"]["wait_frac"], "source": "the same file's SQ passes"}
    if not args.no_cpu:
        result["cpu_baseline"] = uneven64_cpu_baseline(specs, table, keep, args.cpu_budget)
    return result


def uneven64_cpu_baseline(specs, table, keep, budget_s):
    """The reference's per-map loop (numpy_calc_range_angles + eval_sensor_model,
    RangeLib.h:482-555) over the 64 maps of the batch, on the compiled
    reference: every map's steps split over all host threads (median of 3 runs
    per map), and on one thread for every eighth map; the distance transforms
    are built beforehand and timed apart."""
    from oracle import oracle as O
    from range_libc_amd import workloads as W

    if not O.ref_available():
        return None
    threads = len(os.sched_getaffinity(0))
    t_mt = 0.0
    t_1, rays_1 = 0.0, 0
    build_s = 0.0
    total = 0
    for i, (sp, k) in enumerate(zip(specs, keep)):
        img = W.make_map(sp["kind"], sp["side"], sp["map_seed"])
        ins, ang, obs = (t.cpu().numpy() for t in k[1:4])
        t0 = time.perf_counter()
        ref = O.Reference(img, MAX_RANGE, "rm")
        ref.set_sensor_model(table)
        build_s += time.perf_counter() - t0
        rays = ins.shape[0] * ang.shape[0]
        t_mt += float(np.median(timed_runs(lambda: ref.pf_step(ins, ang, obs, threads), 0.0, min_runs=3,
                                           warmup=1)))
        if i % 8 == 0:
            t_1 += float(np.median(timed_runs(lambda: ref.pf_step(ins, ang, obs, 1), 0.0, min_runs=1, warmup=0)))
            rays_1 += rays
        total += rays
        del ref
    return {"value": total / t_mt, "unit": "rays/s", "cores": threads, "kind": "reference",
            "sample": f"all 64 maps, {total} rays, per-map steps on {threads} threads (median of 3 runs per map): "
                      f"{t_mt:.3f} s per batch; distance transforms built beforehand in {build_s:.1f} s",
            "single_thread": {"value": rays_1 / t_1, "unit": "rays/s", "cores": 1, "kind": "reference",
                              "sample": f"every 8th map ({rays_1} rays) on 1 thread: {t_1:.3f} s"},
            "edt_build_s_all_maps": build_s, **host_info()}


def bench_surrogate(args):
    """Secondary workload: the learned surrogate on the particle-filter step's
    map (2048x2048 corridors, max_range 500).  Trains on ``--sg-pairs`` rays
    marched on the GPU (uniform free-space origins and headings, minibatch
    16384; learning rate 1e-3, then 3e-4 and 1e-4 over the last half and
    fifth of the pairs), then casts the step's 4000 x 60 rays through the
    checked path at several bound limits and compares them with the exact
    marcher: error distribution, share of rays finished from a certified start
    point, map reads and dependent read rounds per ray, and time per step (at
    the 0.5-cell bound) against the plain marching kernel."""
    import torch

    import range_libc_amd as rl

    dev = torch.device("cuda", 0)
    stream = torch.cuda.Stream(device=dev)
    torch.cuda.set_stream(stream)
    ctx = rl.default_context(0)
    ctx.set_stream(stream.cuda_stream)
    img, ins, angles, table, true_pose, seed_noise = build_workload()
    m = rl.DeviceMap(ctx, np.ascontiguousarray(img.T.astype(np.uint8)))
    sg = rl.Surrogate(m, seed=0)
    batch = 16384
    t0 = time.perf_counter()
    stages = ((1e-3, 0.5), (3e-4, 0.3), (1e-4, 0.2))
    curve = np.concatenate([sg.train(MAX_RANGE, n_pairs=int(args.sg_pairs * frac), batch=batch, lr=lr, seed=1 + i)
                            for i, (lr, frac) in enumerate(stages)])
    train_s = time.perf_counter() - t0
    n = PARTICLES * BEAMS
    d_ins = torch.from_numpy(ins).to(dev)
    d_ang = torch.from_numpy(angles).to(dev)
    exact = torch.empty(n, dtype=torch.float32, device=dev)
    got = torch.empty(n, dtype=torch.float32, device=dev)
    cnt = torch.empty(n, dtype=torch.int32, device=dev)
    m.numpy_calc_range_angles_dev(rl.RL_METHOD_RM, MAX_RANGE, d_ins.data_ptr(), d_ang.data_ptr(),
                                  exact.data_ptr(), PARTICLES, BEAMS)
    m.count_steps_dev(rl.RL_METHOD_RM, rl.RL_FRAME_ANGLES, MAX_RANGE, d_ins.data_ptr(), d_ang.data_ptr(),
                      cnt.data_ptr(), PARTICLES, BEAMS)
    stream.synchronize()
    ex = exact.cpu().numpy().astype(np.float64)
    exact_reads = float(cnt.sum().item()) / n
    rows = []
    for bmax in (0.5, 2.0, 4.0, 8.0, 16.0, 1e30):
        st = sg.calc_range_angles_dev(MAX_RANGE, d_ins.data_ptr(), d_ang.data_ptr(), got.data_ptr(), PARTICLES,
                                      BEAMS, bound_max=bmax)
        err = np.abs(got.cpu().numpy().astype(np.float64) - ex)
        q = np.quantile(err, [0.5, 0.9, 0.99])
        rows.append({"bound_max": bmax, "from_start": st["from_start"] / n,
                     "fallback_bound": st["fallback_bound"] / n, "fallback_check": st["fallback_check"] / n,
                     "reads_per_ray": st["reads"] / n, "rounds_per_ray": st["rounds"] / n,
                     "err_p50": q[0], "err_p90": q[1], "err_p99": q[2],
                     "err_max": float(err.max()), "exact_share": float(np.mean(err == 0.0))})

    def timed(fn, reps):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        fn()
        torch.cuda._sleep(50_000_000)
        e0.record(stream)
        for _ in range(reps):
            fn()
        e1.record(stream)
        stream.synchronize()
        return e0.elapsed_time(e1) / reps

    plain_ms = timed(lambda: m.numpy_calc_range_angles_dev(rl.RL_METHOD_RM, MAX_RANGE, d_ins.data_ptr(),
                                                           d_ang.data_ptr(), exact.data_ptr(), PARTICLES, BEAMS),
                     args.steps)
    sg_ms = timed(lambda: sg.calc_range_angles_dev(MAX_RANGE, d_ins.data_ptr(), d_ang.data_ptr(), got.data_ptr(),
                                                   PARTICLES, BEAMS, bound_max=0.5, stats=False), args.steps)
    flops = 2.0 * n * (64 * 256 + 2 * 256 * 256 + 256 * 2)
    return {
        "metric": "surrogate rays/s relative to the plain GPU marcher (particle-filter step rays)",
        "value": plain_ms / sg_ms, "unit": "x", "surrogate_ms_per_step": sg_ms, "plain_ms_per_step": plain_ms,
        "surrogate_rays_per_s": n / (sg_ms * 1e-3), "plain_rays_per_s": n / (plain_ms * 1e-3),
        "network_tflops_per_step": flops / 1e12, "network_tflops_s_over_whole_step": flops / (sg_ms * 1e-3) / 1e12,
        "train_pairs": int(curve.size * batch), "train_s": train_s,
        "train_pairs_per_s": curve.size * batch / train_s,
        "loss_first": float(np.mean(curve[:4])), "loss_last": float(np.mean(curve[-16:])),
        "exact_reads_per_ray": exact_reads, "by_bound": rows, "timed_bound_max": 0.5,
        "config": {"workload": "surrogate_corridor2048_p4000_b60_r500", "batch": batch}}


def bench_glt(args):
    """Secondary workload: GiantLUTCast on the particle-filter step's map
    (2048x2048 corridors, max_range 500, 108 heading bins as in the
    reference's test script): the table built on the GPU (one marched ray per
    entry, 453 M rays), the reference's CPU constructor timed on a bounded
    sample (a 256x256 map of the same generator), and the particle-filter step
    answered from the table."""
    import torch

    import range_libc_amd as rl
    from range_libc_amd import workloads as W

    TD = 108
    dev = torch.device("cuda", 0)
    stream = torch.cuda.Stream(device=dev)
    torch.cuda.set_stream(stream)
    ctx = rl.default_context(0)
    ctx.set_stream(stream.cuda_stream)
    img, ins, angles, table, true_pose, seed_noise = build_workload()
    m = rl.DeviceMap(ctx, np.ascontiguousarray(img.T.astype(np.uint8)))
    m.set_sensor_table(table)
    m.build_giant_lut(MAX_RANGE, TD)  # first build: allocation
    builds = []
    for _ in range(3):
        m.build_giant_lut(MAX_RANGE, TD)
        builds.append(m.giant_lut(copy=False)[3])
    build_ms = float(np.median(builds))
    entries = MAP_SIDE * MAP_SIDE * TD
    d_ins = torch.from_numpy(ins).to(dev)
    d_ang = torch.from_numpy(angles).to(dev)
    d_obs = torch.from_numpy(W.noisy_observation(np.full(BEAMS, 50.0, np.float32), seed_noise,
                                                 max_range=MAX_RANGE)).to(dev)
    d_rng = torch.empty(PARTICLES * BEAMS, dtype=torch.float32, device=dev)
    d_w = torch.empty(PARTICLES, dtype=torch.float64, device=dev)

    def step(method):
        m.pf_step_dev(method, MAX_RANGE, d_ins.data_ptr(), d_ang.data_ptr(), d_obs.data_ptr(), d_rng.data_ptr(),
                      d_w.data_ptr(), PARTICLES, BEAMS)

    def timed(fn, reps):
        fn()
        torch.cuda._sleep(50_000_000)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(reps):
            fn()
        e1.record(stream)
        stream.synchronize()
        return e0.elapsed_time(e1) / reps

    glt_ms = timed(lambda: step(rl.RL_METHOD_GLT), args.steps)
    rm_ms = timed(lambda: step(rl.RL_METHOD_RM), args.steps)
    result = {
        "metric": "GiantLUTCast table entries built per second on 1 GPU (2048^2 corridors, 108 bins, R=500)",
        "value": entries / (build_ms * 1e-3), "unit": "entries/s", "build_ms": build_ms, "entries": entries,
        "table_bytes": entries * 2, "pf_step_glt_ms": glt_ms, "pf_step_rm_ms": rm_ms,
        "pf_step_glt_rays_per_s": PARTICLES * BEAMS / (glt_ms * 1e-3),
        "config": {"workload": "glt_corridor2048_td108_r500"}}
    if not args.no_cpu:
        from oracle import oracle as O

        if O.ref_available():
            small = W.corridor_map(256, 1)
            t0 = time.perf_counter()
            O.Reference(small, MAX_RANGE, "glt", theta_disc=TD)
            dt = time.perf_counter() - t0
            n_small = 256 * 256 * TD
            result["cpu_baseline"] = {"value": n_small / dt, "unit": "entries/s", "cores": 1, "kind": "reference",
                                      "sample": f"GiantLUTCast constructor on a 256x256 corridor map, {TD} bins: "
                                                f"{n_small} entries in {dt:.2f} s",
                                      "full_table_estimate_s": entries / (n_small / dt)}
    return result


def bench_bresenham(args):
    """Secondary workload: Bresenham calc_range against ray marching on the
    particle-filter step's map (2048x2048 corridors): 1 M queries from integer
    free cells, headings uniform in [0, 2 pi), max_range 1000, through
    numpy_calc_range (world frame, identity parameters).  Both methods are
    timed on the same device-resident queries; Bresenham's algorithmic bytes
    are one occupancy byte per visited cell (counted on the device) plus 16 B of
    pose and range per ray.  A sample of rays is checked against the oracle and
    every mismatch counted; the reference's CPU path is timed on a sample of
    the same queries over the host's cores."""
    import torch

    import range_libc_amd as rl
    from range_libc_amd import workloads as W

    NQ, R = 1_000_000, 1000.0
    dev = torch.device("cuda", 0)
    stream = torch.cuda.Stream(device=dev)
    torch.cuda.set_stream(stream)
    ctx = rl.default_context(0)
    ctx.set_stream(stream.cuda_stream)
    img = W.corridor_map(MAP_SIDE, 1)
    q = W.world_poses(img, NQ, seed=7, integer=True)
    m = rl.DeviceMap(ctx, np.ascontiguousarray(img.T.astype(np.uint8)))
    d_q = torch.from_numpy(q).to(dev)
    outs = {k: torch.empty(NQ, dtype=torch.float32, device=dev) for k in ("bl", "rm")}
    meth = {"bl": rl.RL_METHOD_BL, "rm": rl.RL_METHOD_RM}
    cells = {}
    for k in meth:
        cnt = torch.zeros(NQ, dtype=torch.int32, device=dev)
        m.count_steps_dev(meth[k], rl.RL_FRAME_WORLD, R, d_q.data_ptr(), 0, cnt.data_ptr(), NQ)
        stream.synchronize()
        cells[k] = int(cnt.sum().item())

    def timed(k, reps):
        fn = lambda: m.numpy_calc_range_dev(meth[k], R, d_q.data_ptr(), outs[k].data_ptr(), NQ)  # noqa: E731
        fn()
        torch.cuda._sleep(50_000_000)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(reps):
            fn()
        e1.record(stream)
        stream.synchronize()
        return e0.elapsed_time(e1) / reps

    reps = max(5, args.steps // 10)
    ms = {k: timed(k, reps) for k in meth}
    # the default particle-filter step (same map; its particles, beams, max_range and sensor model)
    # through either method
    _, ins, angles, table, _, seed_noise = build_workload()
    m.set_sensor_table(table)
    d_ins, d_ang = torch.from_numpy(ins).to(dev), torch.from_numpy(angles).to(dev)
    d_obs = torch.from_numpy(W.noisy_observation(np.full(BEAMS, 50.0, np.float32), seed_noise,
                                                 max_range=MAX_RANGE)).to(dev)
    d_rng = torch.empty(PARTICLES * BEAMS, dtype=torch.float32, device=dev)
    d_w = torch.empty(PARTICLES, dtype=torch.float64, device=dev)

    def pf_timed(method, reps):
        def fn():
            m.pf_step_dev(method, MAX_RANGE, d_ins.data_ptr(), d_ang.data_ptr(), d_obs.data_ptr(),
                          d_rng.data_ptr(), d_w.data_ptr(), PARTICLES, BEAMS)
        fn()
        torch.cuda._sleep(50_000_000)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(reps):
            fn()
        e1.record(stream)
        stream.synchronize()
        return e0.elapsed_time(e1) / reps

    pf_ms = {k: pf_timed(meth[k], max(20, args.steps)) for k in meth}
    alg = {"bl": cells["bl"] + 16 * NQ, "rm": 4 * cells["rm"] + 16 * NQ}
    result = {
        "metric": "Bresenham rays/s on 1 GPU (1 M integer-cell queries, 2048^2 corridors, R=1000)",
        "value": NQ / (ms["bl"] * 1e-3), "unit": "rays/s", "higher_is_better": True,
        "bl_ms": ms["bl"], "rm_ms": ms["rm"], "rm_rays_per_s": NQ / (ms["rm"] * 1e-3),
        "bl_cells_per_ray": cells["bl"] / NQ, "rm_reads_per_ray": cells["rm"] / NQ,
        "pf_step_bl_ms": pf_ms["bl"], "pf_step_rm_ms": pf_ms["rm"],
        "roofline": {"bound": "hbm", "achieved": alg["bl"] / (ms["bl"] * 1e-3) / 1e9, "peak": 8000.0,
                     "unit": "GB/s", "frac": alg["bl"] / (ms["bl"] * 1e-3) / 1e9 / 8000.0,
                     "kernel": "k_cast_bl<world> (occupancy bitmap, walks regrouped)", "alg_bytes_per_launch": alg["bl"],
                     "visited_cells_per_s": cells["bl"] / (ms["bl"] * 1e-3)},
        "config": {"workload": "bresenham_vs_rm_corridor2048_1m_r1000"}}
    # the instruction-issue bounds, from the committed counter pass over this launch
    ipath = os.path.join(ROOT, "profiles", "bresenham_issue_corridor2048_1m.json")
    if os.path.exists(ipath):
        with open(ipath) as f:
            ij = json.load(f)["bresenham"]
        result["issue_bound"] = {
            "valu_us": ij["valu_issue_bound_us"], "salu_us": ij["salu_issue_bound_us"],
            "valu_busy_frac_at_profiled_duration": ij["valu_busy_frac"],
            "valu_busy_frac_here": ij["valu_issue_bound_us"] / (ms["bl"] * 1e3),
            "ta_busy_frac": ij.get("ta_busy_frac"), "l2_requests": ij.get("l2_requests"),
            "source": "profiles/bresenham_issue_corridor2048_1m.json: rocprofv3 --pmc passes (SQ_INSTS_VALU, "
                      "SQ_INSTS_SALU, GRBM_GUI_ACTIVE, TA_BUSY_avr, TCC_REQ_sum) over this launch; not measured "
                      "in this run"}
    from oracle import oracle as O

    pick = np.random.default_rng(0).choice(NQ, 4000, replace=False)
    for k in meth:
        want = O.Oracle(img, R, k).numpy_calc_range(q[pick])
        got = outs[k].cpu().numpy()[pick]
        result[f"{k}_sample_mismatches"] = int(np.sum(got.view(np.uint32) != want.view(np.uint32)))
    result["sample_rays_checked"] = int(pick.size)
    if not args.no_cpu and O.ref_available():
        threads = len(os.sched_getaffinity(0))
        ref = O.Reference(img, R, "bl")
        ref.numpy_calc_range(q[:1000], threads)
        reps, t0 = 0, time.perf_counter()
        while time.perf_counter() - t0 < min(args.cpu_budget, 5.0) or reps == 0:
            ref.numpy_calc_range(q, threads)
            reps += 1
        dt = time.perf_counter() - t0
        result["cpu_baseline"] = {"value": reps * NQ / dt, "unit": "rays/s", "cores": threads,
                                  "kind": "reference",
                                  "sample": f"BresenhamsLine numpy_calc_range over all {NQ} queries, {reps} passes "
                                            f"in {dt:.3f} s"}
    return result


def bench_cddt(args):
    """Secondary workload: CDDTCast / PCDDT on the particle-filter step's map
    (2048x2048 corridors, max_range 500, 108 heading bins): tables built and
    pruned on the GPU, the reference's CPU constructor and prune timed on a
    bounded sample (a 512x512 corridor map), and the particle-filter step
    answered from the pruned tables."""
    import torch

    import range_libc_amd as rl
    from range_libc_amd import workloads as W

    TD = 108
    dev = torch.device("cuda", 0)
    stream = torch.cuda.Stream(device=dev)
    torch.cuda.set_stream(stream)
    ctx = rl.default_context(0)
    ctx.set_stream(stream.cuda_stream)
    img, ins, angles, table, true_pose, seed_noise = build_workload()
    m = rl.DeviceMap(ctx, np.ascontiguousarray(img.T.astype(np.uint8)))
    m.set_sensor_table(table)
    builds, prunes = [], []
    for _ in range(3):
        m.build_cddt(TD)
        builds.append(m.cddt_tables()["build_ms"])
        t0 = time.perf_counter()
        m.prune_cddt(MAX_RANGE)
        prunes.append((time.perf_counter() - t0) * 1e3)
    t = m.cddt_tables()
    d_ins = torch.from_numpy(ins).to(dev)
    d_ang = torch.from_numpy(angles).to(dev)
    d_obs = torch.from_numpy(W.noisy_observation(np.full(BEAMS, 50.0, np.float32), seed_noise,
                                                 max_range=MAX_RANGE)).to(dev)
    d_rng = torch.empty(PARTICLES * BEAMS, dtype=torch.float32, device=dev)
    d_w = torch.empty(PARTICLES, dtype=torch.float64, device=dev)

    def step(method):
        m.pf_step_dev(method, MAX_RANGE, d_ins.data_ptr(), d_ang.data_ptr(), d_obs.data_ptr(), d_rng.data_ptr(),
                      d_w.data_ptr(), PARTICLES, BEAMS)

    def timed(fn, reps):
        fn()
        torch.cuda._sleep(50_000_000)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(reps):
            fn()
        e1.record(stream)
        stream.synchronize()
        return e0.elapsed_time(e1) / reps

    cddt_ms = timed(lambda: step(rl.RL_METHOD_CDDT), args.steps)
    rm_ms = timed(lambda: step(rl.RL_METHOD_RM), args.steps)
    result = {
        "metric": "PCDDT tables built and pruned on 1 GPU (2048^2 corridors, 108 bins, R=500), ms",
        "value": float(np.median(builds)) + float(np.median(prunes)), "unit": "ms", "higher_is_better": False,
        "build_ms": float(np.median(builds)), "prune_ms_wall": float(np.median(prunes)),
        "pruned_entries": int(len(t["vals"])), "bins": int(len(t["start"]) - 1),
        "pf_step_pcddt_ms": cddt_ms, "pf_step_rm_ms": rm_ms,
        "pf_step_pcddt_rays_per_s": PARTICLES * BEAMS / (cddt_ms * 1e-3),
        "config": {"workload": "pcddt_corridor2048_td108_r500"}}
    if not args.no_cpu:
        from oracle import oracle as O

        if O.ref_available():
            small = W.corridor_map(512, 1)
            t0 = time.perf_counter()
            ref = O.Reference(small, MAX_RANGE, "cddt", theta_disc=TD)
            t1 = time.perf_counter()
            ref.cddt_prune(MAX_RANGE)
            t2 = time.perf_counter()
            result["cpu_baseline"] = {"value": (t2 - t0) * 1e3, "unit": "ms", "cores": 1, "kind": "reference",
                                      "sample": f"CDDTCast constructor ({t1 - t0:.2f} s) + prune ({t2 - t1:.2f} s) "
                                                f"on a 512x512 corridor map, {TD} bins (1/16 of the cells)"}
    return result


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU-baseline leg")
    ap.add_argument("--no-secondary", action="store_true",
                    help="default workload: skip the other configurations' figures after the headline")
    ap.add_argument("--cpu-budget", type=float, default=10.0)
    ap.add_argument("--xcd", type=int, default=0, help="XCD-aware particle scheduling (1) or input order (0)")
    ap.add_argument("--box", type=int, default=0, help="stage the map around each particle in LDS (1) or not (0)")
    ap.add_argument("--batch-config", default="", metavar="RPW,BOX,ORDER,GRID",
                    help="uneven64: this batch configuration instead of the autotuner's")
    ap.add_argument("--workload", default="pf_step",
                    choices=["pf_step", "c1", "uneven64", "surrogate", "glt", "cddt", "bresenham"],
                    help="pf_step (the headline particle-filter step), uneven64 (multi-map batch) or "
                         "surrogate (learned start point on the step's map)")
    ap.add_argument("--sg-pairs", type=int, default=50_000_000, help="surrogate training pairs")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run 10 plain device copies of MIB MiB (HBM counter calibration) and exit")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build runs on one GPU only; --gpus must be 1", file=sys.stderr)
        sys.exit(2)

    import torch

    import range_libc_amd as rl

    dev = torch.device("cuda", 0)
    legs = {"c1": bench_c1, "uneven64": bench_uneven64, "surrogate": bench_surrogate, "glt": bench_glt,
            "cddt": bench_cddt, "bresenham": bench_bresenham}
    if args.workload in legs:
        print(json.dumps(legs[args.workload](args)))
        return
    if args.copy_calibration:
        n = args.copy_calibration * (1 << 20) // 4
        src = torch.rand(n, device=dev)
        dst = torch.empty_like(src)
        for _ in range(10):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes_each_way": n * 4, "copies": 10}))
        return
    img, ins, angles, table, true_pose, seed_noise = build_workload()
    from range_libc_amd import workloads as W

    ctx = rl.default_context(0)
    ctx.set_option(rl._lib.RL_OPT_XCD_SCHED, args.xcd)
    ctx.set_option(rl._lib.RL_OPT_LDS_BOX, args.box)
    # a dedicated stream (default priority): torch's current stream is the null
    # stream, which the library would not order its work on
    stream = torch.cuda.Stream(device=dev)
    ctx.set_stream(stream.cuda_stream)
    torch.cuda.set_stream(stream)

    occ = np.ascontiguousarray(img.T.astype(np.uint8))
    # a small map first, so that the timed build does not include loading the
    # EDT kernels' code on their first launch
    rl.DeviceMap(ctx, occ[:64, :64].copy())
    m = rl.DeviceMap(ctx, occ)
    edt_ms = m.edt_build_ms()
    m.set_sensor_table(table)

    d_ins = torch.from_numpy(ins).to(dev)
    d_ang = torch.from_numpy(angles).to(dev)
    d_rng = torch.empty(PARTICLES * BEAMS, dtype=torch.float32, device=dev)
    d_w = torch.empty(PARTICLES, dtype=torch.float64, device=dev)
    # observation: the GPU's own ranges from a seeded true pose plus noise
    d_true = torch.from_numpy(true_pose).to(dev)
    d_tr = torch.empty(BEAMS, dtype=torch.float32, device=dev)
    m.numpy_calc_range_angles_dev(rl.RL_METHOD_RM, MAX_RANGE, d_true.data_ptr(), d_ang.data_ptr(),
                                  d_tr.data_ptr(), 1, BEAMS)
    stream.synchronize()
    obs = W.noisy_observation(d_tr.cpu().numpy(), seed_noise, max_range=MAX_RANGE)
    d_obs = torch.from_numpy(obs).to(dev)

    # algorithmic bytes of one ray-marching launch: per-ray distance-transform reads
    d_cnt = torch.empty(PARTICLES * BEAMS, dtype=torch.int32, device=dev)
    m.count_steps_dev(rl.RL_METHOD_RM, rl.RL_FRAME_ANGLES, MAX_RANGE, d_ins.data_ptr(), d_ang.data_ptr(),
                      d_cnt.data_ptr(), PARTICLES, BEAMS)
    stream.synchronize()
    reads = int(d_cnt.sum().item())
    rays = PARTICLES * BEAMS
    # 4 B per distance-transform read, pose/angle/observation input, range output,
    # one 8-B table entry per ray, one 8-B weight per particle
    alg_bytes = 4 * reads + 12 * PARTICLES + 8 * BEAMS + 4 * rays + 8 * rays + 8 * PARTICLES

    def cast():  # the whole step: one fused launch
        m.pf_step_dev(rl.RL_METHOD_RM, MAX_RANGE, d_ins.data_ptr(), d_ang.data_ptr(), d_obs.data_ptr(),
                      d_rng.data_ptr(), d_w.data_ptr(), PARTICLES, BEAMS)

    def two_launch_step():
        m.numpy_calc_range_angles_dev(rl.RL_METHOD_RM, MAX_RANGE, d_ins.data_ptr(), d_ang.data_ptr(),
                                      d_rng.data_ptr(), PARTICLES, BEAMS)
        m.eval_sensor_model_dev(d_obs.data_ptr(), d_rng.data_ptr(), d_w.data_ptr(), BEAMS, PARTICLES)

    step = cast

    def timed(fn, reps):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(reps):
            fn()
        e1.record(stream)
        stream.synchronize()
        return e0.elapsed_time(e1)

    for _ in range(args.warmup):
        two_launch_step()
    stream.synchronize()
    w_eager = d_w.cpu().numpy().copy()
    r_eager = d_rng.cpu().numpy().copy()
    for _ in range(args.warmup):
        step()
    stream.synchronize()
    assert np.array_equal(d_w.cpu().numpy(), w_eager), "fused step differs from the two calls"
    assert np.array_equal(d_rng.cpu().numpy(), r_eager), "fused step differs from the two calls"
    two_launch_ms = timed(two_launch_step, args.steps) / args.steps

    # (1) eager: one library call per launch from Python
    eager_ms = timed(step, args.steps) / args.steps

    # (2) steady state: GROUP steps captured in one HIP graph, replayed
    GROUP = 10
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph, stream=stream):
        for _ in range(GROUP):
            step()
    for _ in range(3):
        graph.replay()
    stream.synchronize()
    assert np.array_equal(d_w.cpu().numpy(), w_eager), "graph replay changed the weights"
    replays = max(1, args.steps // GROUP)
    t_start = time.perf_counter()
    graph_ms = timed(graph.replay, replays) / (replays * GROUP)
    wall = time.perf_counter() - t_start
    assert np.array_equal(d_w.cpu().numpy(), w_eager), "graph replay changed the weights"

    # (3) the dominant kernel alone, on its own stream: a spin kernel first
    # keeps the queue full so that no host gap enters; one event pair brackets
    # KREPS back-to-back launches (an event record per launch would add its
    # own barrier packet to every bracket; that figure is reported beside it)
    KREPS = 100
    torch.cuda._sleep(100_000_000)
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record(stream)
    for _ in range(KREPS):
        cast()
    b.record(stream)
    stream.synchronize()
    kern_ms = a.elapsed_time(b) / KREPS
    torch.cuda._sleep(100_000_000)
    pairs = []
    for _ in range(KREPS):
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record(stream)
        cast()
        b.record(stream)
        pairs.append((a, b))
    stream.synchronize()
    kern_pair_ms = float(np.mean([a.elapsed_time(b) for a, b in pairs]))
    achieved = alg_bytes / (kern_ms * 1e-3) / 1e9
    step_ms = graph_ms

    # (4) the host-buffer API (pageable numpy in and out, as the reference's
    # Python wrapper is called): the PCIe-inclusive rate, never `value`
    h_rng = np.empty(rays, np.float32)
    h_w = np.empty(PARTICLES, np.float64)
    h_wf = np.empty(PARTICLES, np.float64)

    def host_two_calls():
        m.numpy_calc_range_angles(rl.RL_METHOD_RM, MAX_RANGE, ins, angles, h_rng, PARTICLES, BEAMS)
        m.eval_sensor_model(obs, h_rng, h_w, BEAMS, PARTICLES)

    def host_fused():
        m.repeat_angles_eval_sensor_model(rl.RL_METHOD_RM, MAX_RANGE, ins, angles, obs, h_wf, PARTICLES, BEAMS)

    host = {}
    for name, fn in (("numpy_calc_range_angles+eval_sensor_model", host_two_calls),
                     ("calc_range_repeat_angles_eval_sensor_model", host_fused)):
        for _ in range(5):
            fn()
        reps = max(20, args.steps // 4)
        t0 = time.perf_counter()
        for _ in range(reps):
            fn()
        host[name] = (time.perf_counter() - t0) * 1e3 / reps
    assert np.array_equal(h_w, w_eager) and np.array_equal(h_rng, r_eager), "host API differs from the device step"

    traffic, traffic_src = None, None
    tpath = os.path.join(ROOT, "profiles", "traffic_rm_angles_pf_step.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            tj = json.load(f)
        traffic = tj.get("hbm_bytes_per_launch")
        traffic_src = (f"profiles/traffic_rm_angles_pf_step.json: rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE "
                       f"passes over this kernel (profiled launches {tj.get('profiled_launch_us', '?')} us), "
                       f"scaled by a 256 MiB copy's counts; not measured in this run")

    # the longest dependent chain against the measured latency of one dependent load
    lat = dependent_load_ns(torch, ctx, dev)
    cnt_h = d_cnt.cpu().numpy()
    chain_rounds = longest_chain_rounds(m, ins, angles, cnt_h, ctx, torch, dev)
    chain = {"max_reads_per_ray": int(cnt_h.max()), "longest_chain_rounds": chain_rounds,
             "dependent_load_ns": lat,
             "chain_bound_us_l2": chain_rounds * lat["l2_2MiB"] / 1e3,
             "chain_bound_us_mall": chain_rounds * lat["mall_64MiB"] / 1e3,
             "chain_bound_us_hbm": chain_rounds * lat["hbm_2GiB"] / 1e3}

    result = {
        "metric": "rays/s on 1 GPU (particle-filter step: ray marching over repeated angles + sensor model)",
        "value": rays / (step_ms * 1e-3),
        "unit": "rays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": step_ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "pf_step_corridor2048_4000x60_r500", "map": f"{MAP_SIDE}x{MAP_SIDE} corridor",
                   "particles": PARTICLES, "beams": BEAMS, "max_range": MAX_RANGE, "sensor_table": TABLE,
                   "parallelism": "single_gpu", "xcd_scheduling": bool(args.xcd), "lds_box": bool(args.box)},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "traffic_source": traffic_src,
                     "kernel": "k_particles<RM, sensor> (rl_pf_step_dev)", "kernel_ms": kern_ms,
                     "kernel_ms_event_pair_per_launch": kern_pair_ms, "kernel_launches_timed": KREPS,
                     "alg_bytes_per_launch": alg_bytes, "edt_reads_per_ray": reads / rays},
        "chain": chain,
        "eager_ms_per_step": eager_ms,
        "two_launch_eager_ms_per_step": two_launch_ms,
        "graph_group": GROUP,
        "edt_build_ms": edt_ms,
        "host_api_ms_per_step": {k: round(v, 4) for k, v in host.items()},
        "host_api_rays_per_s": {k: rays / (v * 1e-3) for k, v in host.items()},
        "wall_s_timed": wall,
    }
    if not args.no_cpu:
        result["cpu_baseline"] = cpu_baseline(img, ins, angles, obs, args.cpu_budget)
    if not args.no_secondary:
        result["secondary"] = secondary_workloads(args, legs)
    print(json.dumps(result))


def secondary_workloads(args, legs) -> dict:
    """The other configurations, measured in the same run after the headline
    (device time only, fewer timed launches, no CPU legs): the key figures of
    each leg's own JSON line (`bench.py --workload NAME` prints it in full)."""
    sub = argparse.Namespace(**vars(args))
    sub.steps, sub.warmup, sub.no_cpu = max(20, args.steps // 10), 5, True
    keys = ("metric", "value", "unit", "ms_per_step", "bl_ms", "rm_ms", "rays", "mismatches_vs_oracle",
            "bl_sample_mismatches", "rm_sample_mismatches", "pf_step_bl_ms", "build_ms", "prune_ms_wall", "pf_step_glt_ms",
            "pf_step_pcddt_ms", "pf_step_rm_ms", "surrogate_ms_per_step", "plain_ms_per_step", "train_s",
            "autotuned")
    out = {}
    for name in ("c1", "bresenham", "uneven64", "glt", "cddt", "surrogate"):
        t0 = time.perf_counter()
        try:
            r = legs[name](sub)
            d = {k: r[k] for k in keys if k in r}
            if "roofline" in r:
                d["roofline_frac"] = r["roofline"].get("frac")
            if name == "surrogate":
                d["from_start_at_0.5"] = r["by_bound"][0]["from_start"]
                d["err_p99_at_0.5"] = r["by_bound"][0]["err_p99"]
            out[name] = d
        except Exception as e:  # a failing leg is reported, not fatal to the headline line
            out[name] = {"error": f"{type(e).__name__}: {e}"}
        out[name]["leg_wall_s"] = round(time.perf_counter() - t0, 2)
    return out


if __name__ == "__main__":
    main()
