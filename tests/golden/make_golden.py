"""Write the golden vectors under tests/golden/ from the compiled reference.

Run in a container that has the reference source tree (it builds
oracle/_ref/libref.so from it, see oracle/Makefile):

    make -C oracle ref && python tests/golden/make_golden.py

Every expected output here comes from the reference's own C++
(includes/RangeLib.h, vendor/distance_transform.h) compiled with parity flags;
inputs are seeded.  The PNG maps copied into maps/ are data files of the
reference (maps/*.png; pywrapper/test.py loads synthetic.map.png).
"""
from __future__ import annotations

import hashlib
import math
import os
import shutil
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402
from range_libc_amd import workloads as W  # noqa: E402

REF_MAPS = "/root/reference/maps"
PNGS = ["small.map.png", "quad.map.png", "single_pixel.map.png", "synthetic.map.png",
        "test.png", "basement_hallways_5cm.png"]


def sha(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def queries(img, n, seed, frame):
    g = W.grid_poses(img, n, seed)
    return g if frame == "grid" else W.to_world(g)


def edge_queries(w, h, seed):
    """Origins on and outside the map edges, integer and fractional, plus
    axis-aligned and diagonal headings."""
    rng = np.random.default_rng(seed)
    xs = np.concatenate([rng.uniform(-3, w + 3, 200), np.arange(-1, w + 2, dtype=np.float64),
                         np.array([0.0, w - 1e-3, w, 0.5, w - 0.5])])
    ys = np.concatenate([rng.uniform(-3, h + 3, 200), np.arange(-1, w + 2, dtype=np.float64) % (h + 2) - 1,
                         np.array([0.0, h - 1e-3, h, 0.5, h - 0.5])])
    heads = np.array([0, math.pi / 2, math.pi, 3 * math.pi / 2, math.pi / 4, -math.pi / 4, 2 * math.pi,
                      1e-7, -1e-7, 10.0, -10.0])
    n = len(xs)
    th = np.concatenate([heads[np.arange(n) % len(heads)][: n // 2], rng.uniform(-8, 8, n - n // 2)])
    return np.stack([xs, ys, th], 1).astype(np.float32)


def main():
    if not O.ref_available():
        sys.exit("oracle/_ref/libref.so missing: run `make -C oracle ref` first")
    maps_dir = os.path.join(HERE, "maps")
    os.makedirs(maps_dir, exist_ok=True)
    for f in PNGS:
        shutil.copyfile(os.path.join(REF_MAPS, f), os.path.join(maps_dir, f))

    # --- PNG maps: decoded occupancy, distance transform, ranges -------------
    out = {}
    for f in PNGS:
        key = f.replace(".png", "").replace(".", "_")
        for thr in (128.0, 1.0):
            img = O.ref_load_png(os.path.join(maps_dir, f), thr)
            tag = f"{key}_t{int(thr)}"
            out[f"{tag}_occ"] = np.packbits(img.reshape(-1))
            out[f"{tag}_shape"] = np.array(img.shape)
        img = O.ref_load_png(os.path.join(maps_dir, f), 128.0)
        edt = O.ref_edt(img)
        if edt.size <= 4096:
            out[f"{key}_edt"] = edt
        else:
            rng = np.random.default_rng(1)
            idx = rng.integers(0, edt.size, 4096)
            out[f"{key}_edt_idx"] = idx
            out[f"{key}_edt_val"] = edt.reshape(-1)[idx]
            out[f"{key}_edt_sha"] = np.array(sha(edt))
        h, w = img.shape
        eq = edge_queries(w, h, 5)
        for kind in ("rm", "bl"):
            ref = O.Reference(img, 500.0, kind)
            out[f"{key}_{kind}_edge_ins"] = eq
            out[f"{key}_{kind}_edge_grid"] = ref.calc_range_many(eq)
            out[f"{key}_{kind}_edge_world"] = ref.numpy_calc_range(eq)
            if img.size > 64 and (img == 0).any():
                q = queries(img, 2000, 6, "world")
                out[f"{key}_{kind}_q"] = q
                out[f"{key}_{kind}_world"] = ref.numpy_calc_range(q)
    np.savez_compressed(os.path.join(HERE, "png_maps.npz"), **out)

    # --- seeded 512^2 rectangles + disks (the 100k-query workload's map) ------
    img = W.rect_disk_map(512, seed=0)
    g = {"map_sha": np.array(sha(img)), "edt_sha": np.array(sha(O.ref_edt(img)))}
    rng = np.random.default_rng(2)
    idx = rng.integers(0, img.size, 4096)
    g["edt_idx"], g["edt_val"] = idx, O.ref_edt(img).reshape(-1)[idx]
    table = W.sensor_table(501)
    g["table_sha"] = np.array(sha(table))
    angles = W.beam_angles(60)
    world = (0.05, 0.3, -3.0, 2.0, math.sin(0.3), math.cos(0.3))
    g["world"] = np.array(world, dtype=np.float32)
    for kind in ("rm", "bl"):
        ref = O.Reference(img, 500.0, kind)
        ref.set_sensor_model(table)
        qw = queries(img, 3000, 3, "world")
        qg = queries(img, 3000, 4, "grid")
        parts = queries(img, 50, 5, "world")
        g[f"{kind}_qw"], g[f"{kind}_world"] = qw, ref.numpy_calc_range(qw)
        g[f"{kind}_qg"], g[f"{kind}_grid"] = qg, ref.calc_range_many(qg)
        g[f"{kind}_parts"] = parts
        rr = ref.numpy_calc_range_angles(parts, angles)
        g[f"{kind}_angles"] = rr
        obs = W.noisy_observation(rr[:60], seed=6)
        g[f"{kind}_obs"] = obs
        g[f"{kind}_sensor"] = ref.eval_sensor_model(obs, rr, 60, 50)
        g[f"{kind}_fused"] = ref.calc_range_repeat_angles_eval_sensor_model(parts, angles, obs)
        refw = O.Reference(img, 500.0, kind, world)
        c, s = math.cos(0.3), math.sin(0.3)
        gg = qw.astype(np.float64)
        wq = np.stack([(c * gg[:, 0] + s * gg[:, 1]) * 0.05 - 3.0, (-s * gg[:, 0] + c * gg[:, 1]) * 0.05 + 2.0,
                       gg[:, 2] - 1.0], 1).astype(np.float32)
        g[f"{kind}_wq"], g[f"{kind}_wworld"] = wq, refw.numpy_calc_range(wq)
        wparts = wq[:40]
        g[f"{kind}_wangles"] = refw.numpy_calc_range_angles(wparts, angles)
    np.savez_compressed(os.path.join(HERE, "rect_disk_512.npz"), **g)

    # --- 2048^2 corridor map (particle-filter workload) -----------------------
    img = W.corridor_map(2048, seed=1)
    c = {"map_sha": np.array(sha(img))}
    edt = O.ref_edt(img)
    c["edt_sha"] = np.array(sha(edt))
    idx = np.random.default_rng(7).integers(0, img.size, 4096)
    c["edt_idx"], c["edt_val"] = idx, edt.reshape(-1)[idx]
    parts = W.world_poses(img, 200, seed=8)
    c["parts"] = parts
    for kind in ("rm", "bl"):
        ref = O.Reference(img, 500.0, kind)
        c[f"{kind}_angles"] = ref.numpy_calc_range_angles(parts, angles)
        qi = W.world_poses(img, 2000, seed=9, integer=True)
        c[f"{kind}_qi"], c[f"{kind}_r1000"] = qi, O.Reference(img, 1000.0, kind).numpy_calc_range(qi)
    np.savez_compressed(os.path.join(HERE, "corridor_2048.npz"), **c)

    # --- degenerate maps ------------------------------------------------------
    d = {}
    cases = {"empty": np.zeros((17, 23), np.uint8), "full": np.ones((9, 7), np.uint8),
             "one_free": np.pad(np.zeros((1, 1), np.uint8), 2, constant_values=1),
             "checker": (np.indices((12, 12)).sum(0) % 2).astype(np.uint8),
             "tall": (np.random.default_rng(10).random((300, 7)) < 0.02).astype(np.uint8)}
    for name, img in cases.items():
        d[f"{name}_img"] = img
        d[f"{name}_edt"] = O.ref_edt(img)
        h, w = img.shape
        eq = edge_queries(w, h, 11)
        d[f"{name}_ins"] = eq
        for kind in ("rm", "bl"):
            for mr in (0.0, 1.0, 50.0):
                ref = O.Reference(img, mr, kind)
                d[f"{name}_{kind}_{int(mr)}_grid"] = ref.calc_range_many(eq)
                d[f"{name}_{kind}_{int(mr)}_world"] = ref.numpy_calc_range(eq)
    np.savez_compressed(os.path.join(HERE, "degenerate.npz"), **d)

    giant_lut()
    cddt()
    listing()


def glt_cases():
    """(name, image, max_range, theta_discretization) of the GiantLUTCast fixtures."""
    return [("rect64", W.rect_disk_map(64, seed=3), 60.0, 108),
            ("corr96x80", np.ascontiguousarray(W.corridor_map(96, seed=2)[:, :80]), 120.0, 37),
            ("rect48", W.rect_disk_map(48, seed=5), 500.0, 360)]


def glt_queries(img, td, seed):
    """Poses over and around the map; headings: bin centres, [-30, 30), [-3000, 3000)."""
    rng = np.random.default_rng(seed)
    h, w = img.shape
    n = 3000
    heads = np.concatenate([(np.arange(n // 3) % td) * np.float32(2 * math.pi / td), rng.uniform(-30, 30, n // 3),
                            rng.uniform(-3000, 3000, n - 2 * (n // 3))])
    return np.stack([rng.uniform(-5, w + 5, n), rng.uniform(-5, h + 5, n), heads], axis=1).astype(np.float32)


def giant_lut():
    """GiantLUTCast (RangeLib.h:1772-1900) built and queried by the compiled reference."""
    g = {}
    world = (0.05, 0.3, -3.0, 2.0, math.sin(0.3), math.cos(0.3))
    for name, img, R, td in glt_cases():
        ref = O.Reference(img, R, "glt", theta_disc=td)
        table, consts = ref.glt_table()
        g[f"{name}_img"] = img
        g[f"{name}_table_sha"] = np.array(sha(table))
        idx = np.random.default_rng(3).integers(0, table.size, 5000).astype(np.int32)
        g[f"{name}_sample_idx"] = idx
        g[f"{name}_sample"] = table.reshape(-1)[idx]
        g[f"{name}_consts"] = consts
        q = glt_queries(img, td, 4)
        g[f"{name}_ins"] = q
        g[f"{name}_grid"] = ref.calc_range_many(q)
        g[f"{name}_world"] = ref.numpy_calc_range(q)
        ang = W.beam_angles(61)
        g[f"{name}_angles"] = ref.numpy_calc_range_angles(q[:300], ang)
        refw = O.Reference(img, R, "glt", world=world, theta_disc=td)
        g[f"{name}_world_rot"] = refw.numpy_calc_range(q)
    np.savez_compressed(os.path.join(HERE, "giant_lut.npz"), **g)


def cddt_cases():
    """(name, image, max_range, theta_discretization) of the CDDTCast fixtures."""
    return [("rect64", W.rect_disk_map(64, seed=3), 60.0, 108),
            ("corr96x80", np.ascontiguousarray(W.corridor_map(96, seed=2)[:, :80]), 120.0, 37),
            ("corr300", W.corridor_map(300, seed=4), 200.0, 120)]


def cddt():
    """CDDTCast and its pruned form built and queried by the compiled reference:
    table hashes (widths, translations, bin offsets, values) and queries from
    inside the map."""
    g = {}
    world = (0.05, 0.3, -3.0, 2.0, math.sin(0.3), math.cos(0.3))
    for name, img, R, td in cddt_cases():
        h, w = img.shape
        g[f"{name}_img"] = img
        rng = np.random.default_rng(5)
        n = 3000
        q = np.stack([rng.uniform(0, w, n), rng.uniform(0, h, n),
                      np.concatenate([rng.uniform(-30, 30, n // 2), rng.uniform(-3000, 3000, n - n // 2)])],
                     axis=1).astype(np.float32)
        g[f"{name}_ins"] = q
        for tag, prune in (("full", False), ("pruned", True)):
            ref = O.Reference(img, R, "cddt", theta_disc=td)
            if prune:
                ref.cddt_prune(R)
            t = ref.cddt_tables()
            for k in ("widths", "trans", "start", "vals"):
                g[f"{name}_{tag}_{k}_sha"] = np.array(sha(t[k]))
            g[f"{name}_{tag}_nvals"] = np.array(len(t["vals"]))
            g[f"{name}_{tag}_json_sha"] = np.array(hashlib.sha256(ref.cddt_json().encode()).hexdigest())
            g[f"{name}_{tag}_grid"] = ref.calc_range_many(q)
            # calc_range_many_radial_optimized from free-cell world poses, 60 beams over 270 degrees
            g[f"{name}_{tag}_radial"] = ref.calc_range_many_radial_optimized(
                W.world_poses(img, 200, seed=6), 60, -3 * math.pi / 4, 3 * math.pi / 4)
        # world-frame queries land outside the map for this world: keep those inside
        refw = O.Reference(img, R, "cddt", world=world, theta_disc=td)
        g[f"{name}_world_rot"] = refw.numpy_calc_range(world_inside(q, world, w, h))
    np.savez_compressed(os.path.join(HERE, "cddt.npz"), **g)


def world_inside(q, world, w, h):
    """World poses whose grid origin lies inside the map, half a cell from its
    edges.  The reference's transform (RangeLib.h:463-475) rotates
    (x, y) = ((xw - ox) / s, (yw - oy) / s) by the map angle and marches from
    (y', x'); this inverts it for origins (q[:, 0], q[:, 1]) scaled into
    [0.5, w - 0.5) x [0.5, h - 0.5)."""
    scale, ang, ox, oy, s, c = world
    gx = 0.5 + q[:, 0].astype(np.float64) / w * (w - 1.0)  # march origin, first grid index
    gy = 0.5 + q[:, 1].astype(np.float64) / h * (h - 1.0)
    xr, yr = gy, gx  # (x', y') with calc_range(y', x', .)
    x = c * xr + s * yr
    y = -s * xr + c * yr
    out = q.copy()
    out[:, 0] = (x * scale + ox).astype(np.float32)
    out[:, 1] = (y * scale + oy).astype(np.float32)
    return out


def listing():
    for f in sorted(os.listdir(HERE)):
        p = os.path.join(HERE, f)
        if os.path.isfile(p):
            print(f"{f:32s} {os.path.getsize(p):9d} B")


if __name__ == "__main__":
    if sys.argv[1:] == ["glt"]:
        giant_lut()
        listing()
    elif sys.argv[1:] == ["cddt"]:
        cddt()
        listing()
    else:
        main()
