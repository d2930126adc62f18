"""Seeded synthetic workloads for the ray-casting path.

Maps are returned as numpy ``uint8`` images ``img[y, x]`` (1 = occupied), the
layout ``PyOMap(ndarray)`` takes (RangeLibc.pyx:140-145: ``grid[x][y] =
arr[y, x]``).  All randomness comes from ``numpy.random.default_rng(seed)``.

Coordinate frames.  The reference's batched Python entry points
(``calc_range_many``, ``calc_range_repeat_angles``) apply the world-to-grid
conversion (RangeLib.h:463-475), which with identity world parameters casts
from grid point ``(ins[:,1], ins[:,0])`` at heading ``-theta - 3*pi/2``.  So a
pose for those calls is ``(grid_y, grid_x, theta)`` = ``(row, column, theta)``
of the image; ``grid_poses`` returns grid-frame ``(grid_x, grid_y, theta)`` for
the C++ ``calc_range_many`` which does no conversion (RangeLib.h:819-831).
"""
from __future__ import annotations

import math

import numpy as np

__all__ = [
    "rect_disk_map",
    "corridor_map",
    "free_cells",
    "world_poses",
    "grid_poses",
    "to_world",
    "sensor_table",
    "beam_angles",
    "noisy_observation",
    "uneven_batch_specs",
    "make_map",
]


def rect_disk_map(size: int = 512, n_rect: int = 25, n_disk: int = 15, seed: int = 0,
                  border: int = 2, rmin: int = 5, rmax: int = 40) -> np.ndarray:
    """Square map with random axis-aligned rectangles, disks and an outer border.

    With the defaults (512 px, 25 rectangles of side 15-80 px, 15 disks of
    radius 5-40 px, 2-px border) about 30% of the cells are occupied.
    Rectangle sides and counts scale with ``size / 512`` for larger maps.
    """
    rng = np.random.default_rng(seed)
    s = size / 512.0
    img = np.zeros((size, size), dtype=np.uint8)
    img[:border, :] = 1
    img[-border:, :] = 1
    img[:, :border] = 1
    img[:, -border:] = 1
    n_rect = max(1, int(round(n_rect * s * s)))
    n_disk = max(1, int(round(n_disk * s * s)))
    for _ in range(n_rect):
        w = int(rng.integers(15, 81))
        h = int(rng.integers(15, 81))
        x0 = int(rng.integers(0, max(1, size - w)))
        y0 = int(rng.integers(0, max(1, size - h)))
        img[y0:y0 + h, x0:x0 + w] = 1
    for _ in range(n_disk):
        r = float(rng.uniform(rmin, rmax))
        cx = float(rng.uniform(0, size))
        cy = float(rng.uniform(0, size))
        x_lo, x_hi = max(0, int(cx - r) - 1), min(size, int(cx + r) + 2)
        y_lo, y_hi = max(0, int(cy - r) - 1), min(size, int(cy + r) + 2)
        yy, xx = np.ogrid[y_lo:y_hi, x_lo:x_hi]
        sub = (xx + 0.5 - cx) ** 2 + (yy + 0.5 - cy) ** 2
        img[y_lo:y_hi, x_lo:x_hi][sub <= r * r] = 1
    return img


def _wall_layout(size: int, rng, wmin: int, wmax: int, cmin: int, cmax: int):
    """Alternating wall / corridor intervals along one axis, walls at both ends."""
    walls = []
    pos = 0
    while True:
        w = int(rng.integers(wmin, wmax + 1))
        walls.append((pos, min(size, pos + w)))
        pos += w
        c = int(rng.integers(cmin, cmax + 1))
        if pos + c + wmax >= size:
            # close the map with a final wall flush with the edge
            last = int(rng.integers(wmin, wmax + 1))
            walls.append((size - last, size))
            break
        pos += c
    return walls


def corridor_map(size: int = 2048, seed: int = 1, wall_min: int = 3, wall_max: int = 8,
                 corridor_min: int = 20, corridor_max: int = 60, door_prob: float = 0.7,
                 door_min: int = 8, door_max: int = 20) -> np.ndarray:
    """Grid of rooms: walls 3-8 px thick, corridors 20-60 px wide, random doors.

    Vertical and horizontal walls split the square into rooms; each wall
    segment between two crossing walls gets a door gap with probability
    ``door_prob``.
    """
    rng = np.random.default_rng(seed)
    img = np.zeros((size, size), dtype=np.uint8)
    xw = _wall_layout(size, rng, wall_min, wall_max, corridor_min, corridor_max)
    yw = _wall_layout(size, rng, wall_min, wall_max, corridor_min, corridor_max)
    for a, b in xw:
        img[:, a:b] = 1
    for a, b in yw:
        img[a:b, :] = 1
    # doors in vertical walls (between consecutive horizontal walls)
    for i, (a, b) in enumerate(xw):
        if i == 0 or i == len(xw) - 1:
            continue
        for j in range(len(yw) - 1):
            lo, hi = yw[j][1], yw[j + 1][0]
            if hi - lo < door_min + 2 or rng.random() >= door_prob:
                continue
            dw = int(rng.integers(door_min, min(door_max, hi - lo - 1) + 1))
            d0 = int(rng.integers(lo, hi - dw + 1))
            img[d0:d0 + dw, a:b] = 0
    for i, (a, b) in enumerate(yw):
        if i == 0 or i == len(yw) - 1:
            continue
        for j in range(len(xw) - 1):
            lo, hi = xw[j][1], xw[j + 1][0]
            if hi - lo < door_min + 2 or rng.random() >= door_prob:
                continue
            dw = int(rng.integers(door_min, min(door_max, hi - lo - 1) + 1))
            d0 = int(rng.integers(lo, hi - dw + 1))
            img[a:b, d0:d0 + dw] = 0
    return img


def make_map(kind: str, size: int, seed: int) -> np.ndarray:
    if kind == "corridor":
        return corridor_map(size, seed)
    if kind == "rect_disk":
        return rect_disk_map(size, seed=seed)
    raise ValueError(f"unknown map kind {kind!r}")


def free_cells(img: np.ndarray) -> np.ndarray:
    """(n, 2) int array of free cells as (grid_x, grid_y) = (column, row)."""
    ys, xs = np.nonzero(img == 0)
    return np.stack([xs, ys], axis=1)


def grid_poses(img: np.ndarray, n: int, seed: int, integer: bool = False) -> np.ndarray:
    """n float32 grid-frame poses (grid_x, grid_y, theta) in free cells.

    Origins are a uniformly chosen free cell plus a U[0,1)^2 sub-cell offset
    (or the cell's integer corner when ``integer``); theta ~ U[0, 2*pi).
    """
    rng = np.random.default_rng(seed)
    cells = free_cells(img)
    pick = cells[rng.integers(0, len(cells), size=n)]
    off = np.zeros((n, 2)) if integer else rng.random((n, 2))
    theta = rng.random(n) * (2.0 * math.pi)
    out = np.empty((n, 3), dtype=np.float32)
    out[:, 0] = (pick[:, 0] + off[:, 0]).astype(np.float32)
    out[:, 1] = (pick[:, 1] + off[:, 1]).astype(np.float32)
    out[:, 2] = theta.astype(np.float32)
    # a sub-cell offset can round up to the next cell edge in float32
    np.minimum(out[:, 0], np.nextafter(pick[:, 0] + 1.0, 0, dtype=np.float32), out=out[:, 0])
    np.minimum(out[:, 1], np.nextafter(pick[:, 1] + 1.0, 0, dtype=np.float32), out=out[:, 1])
    return out


def to_world(grid: np.ndarray) -> np.ndarray:
    """Grid-frame poses -> the (row, column, theta) layout the world-frame calls
    take under identity world parameters (see module docstring)."""
    out = np.empty_like(grid)
    out[:, 0] = grid[:, 1]
    out[:, 1] = grid[:, 0]
    out[:, 2] = grid[:, 2]
    return out


def world_poses(img: np.ndarray, n: int, seed: int, integer: bool = False) -> np.ndarray:
    """n float32 poses in free cells for the world-frame calls (identity params)."""
    return to_world(grid_poses(img, n, seed, integer))


def beam_angles(n: int = 60, fov: float = 1.5 * math.pi) -> np.ndarray:
    """n float32 beam angles evenly spanning ``fov`` (270 degrees), centred on 0."""
    return np.linspace(-fov / 2.0, fov / 2.0, n).astype(np.float32)


def sensor_table(k: int = 501, sigma_hit: float = 8.0, z_hit: float = 0.75, z_short: float = 0.01,
                 z_max: float = 0.07, z_rand: float = 0.12, lambda_short: float = 0.01) -> np.ndarray:
    """k x k beam model ``table[r, d]`` (r = observed, d = expected range, px).

    Mixture of a Gaussian hit term (sigma 8 px), an exponential short-reading
    term, a max-range spike and a uniform term; each column d is normalised to
    sum to 1.  Built in float32 and returned as float64 (the type
    ``set_sensor_model`` takes, RangeLibc.pyx:336-340).
    """
    r = np.arange(k, dtype=np.float64)[:, None]
    d = np.arange(k, dtype=np.float64)[None, :]
    hit = np.exp(-0.5 * ((r - d) / sigma_hit) ** 2) / (sigma_hit * math.sqrt(2 * math.pi))
    short = np.where(r <= d, lambda_short * np.exp(-lambda_short * r), 0.0)
    mx = np.where(r == k - 1, 1.0, 0.0) * np.ones_like(d)
    rand = np.full((k, k), 1.0 / (k - 1))
    t = z_hit * hit + z_short * short + z_max * mx + z_rand * rand
    t /= t.sum(axis=0, keepdims=True)
    return t.astype(np.float32).astype(np.float64)


def noisy_observation(true_ranges: np.ndarray, seed: int, sigma: float = 2.0,
                      max_range: float = 500.0) -> np.ndarray:
    """Observed ranges = true ranges + N(0, sigma^2), clipped to [0, max_range]."""
    rng = np.random.default_rng(seed)
    obs = true_ranges.astype(np.float64) + rng.normal(0.0, sigma, size=true_ranges.shape)
    return np.clip(obs, 0.0, max_range).astype(np.float32)


def uneven_batch_specs(n_maps: int = 64, seed: int = 4, side_min: int = 256, side_max: int = 4096,
                       p_min: int = 500, p_max: int = 8000, b_min: int = 30, b_max: int = 120):
    """Specs of the uneven multi-map batch: side log-uniform in [256, 4096],
    generator by seed, 500-8000 particles and 30-120 beams per map."""
    rng = np.random.default_rng(seed)
    specs = []
    for i in range(n_maps):
        side = int(round(math.exp(rng.uniform(math.log(side_min), math.log(side_max)))))
        kind = "corridor" if rng.random() < 0.5 else "rect_disk"
        specs.append({
            "side": side,
            "kind": kind,
            "map_seed": int(rng.integers(0, 2**31 - 1)),
            "particles": int(rng.integers(p_min, p_max + 1)),
            "beams": int(rng.integers(b_min, b_max + 1)),
            "pose_seed": int(rng.integers(0, 2**31 - 1)),
        })
    return specs
