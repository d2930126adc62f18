"""Map ingestion feeding the kernels: a ROS OccupancyGrid message (RangeLibc.pyx:146-167)
and a PNG file (RangeLib.h:158-201) loaded through PyOMap, then cast on the GPU in the
map's world frame and compared bit for bit with the oracle given the same occupancy and
world parameters."""
import math
import os
from types import SimpleNamespace as NS

import numpy as np
import pytest

import range_libc_amd as rl
from oracle import oracle as O
from range_libc_amd import workloads as W
from tests.conftest import GOLDEN, assert_bits_equal

pytestmark = pytest.mark.gpu


def _world_poses_from_grid(omap, grid_xy_theta):
    """World-frame poses whose world-to-grid transform (RangeLib.h:442-450)
    lands near the given grid cells: the inverse rotation and scale in double,
    so the grid positions differ from the input by rounding only."""
    scale, angle, ox, oy = omap.world[:4]
    gx, gy, th = grid_xy_theta[:, 0], grid_xy_theta[:, 1], grid_xy_theta[:, 2]
    c, s = math.cos(-angle), math.sin(-angle)
    xw = (c * gx - s * gy) * scale + ox
    yw = (s * gx + c * gy) * scale + oy
    return np.stack([xw, yw, th], 1).astype(np.float32)


def _check_methods(omap, img, n=600, m=40, R=150.0, seed=3):
    rng = np.random.default_rng(seed)
    free = np.argwhere(img == 0)  # (row y, col x) of img[y, x]
    pick = free[rng.integers(0, len(free), n)]
    # numpy_calc_range casts from grid (y', x') = the transformed (x, y): the
    # pose's first coordinate maps to the grid's second index
    grid = np.stack([pick[:, 0] + rng.random(n), pick[:, 1] + rng.random(n), rng.uniform(-np.pi, np.pi, n)], 1)
    ins = _world_poses_from_grid(omap, grid)
    ang = W.beam_angles(m)
    table = W.sensor_table(201)
    for cls, kind in (("PyRayMarchingGPU", "rm"), ("PyBresenhamsLine", "bl")):
        meth = getattr(rl, cls)(omap, R)
        meth.set_sensor_model(table)
        orc = O.Oracle(img, R, kind, world=omap.world)
        orc.set_sensor_model(table)
        out = np.empty(n, np.float32)
        meth.calc_range_many(ins, out)
        assert_bits_equal(out, orc.numpy_calc_range(ins), f"{cls} calc_range_many")
        rr = np.empty(n * m, np.float32)
        meth.calc_range_repeat_angles(ins, ang, rr)
        want = orc.numpy_calc_range_angles(ins, ang)
        assert_bits_equal(rr, want, f"{cls} repeat_angles")
        # many rays hit something before max_range (ranges come out in world units):
        # the poses land inside the map
        assert np.mean(want < np.float32(R * omap.world[0])) > 0.2
        obs = W.noisy_observation(want[:m], seed=seed, max_range=R)
        w = np.empty(n, np.float64)
        meth.calc_range_repeat_angles_eval_sensor_model(ins, ang, obs, w)
        assert np.array_equal(w, orc.calc_range_repeat_angles_eval_sensor_model(ins, ang, obs)), cls


def test_ros_occupancy_grid_feeds_the_kernels():
    img_src = W.corridor_map(256, seed=9)  # the message's rows and columns
    h, w = img_src.shape
    data = np.where(img_src > 0, 100, np.where(np.arange(h * w).reshape(h, w) % 7 == 0, -1, 0)).astype(np.int8)
    yaw = 0.3
    q = NS(x=0.0, y=0.0, z=math.sin(yaw / 2), w=math.cos(yaw / 2))
    msg = NS(info=NS(width=w, height=h, resolution=0.05,
                     origin=NS(position=NS(x=1.5, y=-2.0, z=0.0), orientation=q)),
             data=[int(v) for v in data.ravel()])
    omap = rl.PyOMap(msg)
    assert abs(omap.world[1] + yaw) < 1e-12 and omap.world[0] == 0.05
    img = np.ascontiguousarray(omap.occ.T)  # img[y, x] of the grid the reference builds
    assert np.array_equal(omap.occ, (data.reshape(h, w) > 10).astype(np.uint8))
    _check_methods(omap, img)


def test_png_file_feeds_the_kernels():
    z = np.load(os.path.join(GOLDEN, "png_maps.npz"))
    shape = tuple(int(v) for v in z["basement_hallways_5cm_t128_shape"])
    want_img = np.unpackbits(z["basement_hallways_5cm_t128_occ"])[: shape[0] * shape[1]].reshape(shape)
    omap = rl.PyOMap(os.path.join(GOLDEN, "maps", "basement_hallways_5cm.png"), 128.0)
    assert not omap.error() and np.array_equal(omap.occ.T, want_img)
    omap.set_world(0.05, -0.7, 3.0, 4.0)
    _check_methods(omap, want_img, seed=5)
