"""range_libc_amd: RangeLibc's batched 2-D ray casting on AMD MI355X (gfx950).

Python surface of the reference's Cython module (pywrapper/RangeLibc.pyx):
same class names, constructor arguments, method names, argument order and
array conventions.  Every method runs on the GPU through the C ABI of
``include/rangelib_hip.h`` (librangelib_hip.so); there is no CPU path.

    import range_libc_amd as range_libc
    omap = range_libc.PyOMap(img)                  # img[y, x], nonzero = occupied
    rm = range_libc.PyRayMarchingGPU(omap, 500)
    rm.calc_range_repeat_angles(poses, angles, ranges)
"""
from __future__ import annotations

import math
import sys

import numpy as np

from . import _lib
from ._lib import (RL_FRAME_ANGLES, RL_FRAME_GRID, RL_FRAME_WORLD, RL_METHOD_BL, RL_METHOD_CDDT, RL_METHOD_GLT,
                   RL_METHOD_RM, Batch,
                   Context, DeviceMap, RangeLibError, Surrogate, default_context)

__all__ = [
    "PyOMap", "PyRayMarchingGPU", "PyRayMarching", "PyBresenhamsLine", "PyCDDTCast", "PyGiantLUTCast",
    "PyNull",
    "RangeLibError", "Context", "DeviceMap", "Batch", "Surrogate", "default_context", "cddt_json",
    "USE_CACHED_TRIG", "USE_ALTERNATE_MOD", "USE_CACHED_CONSTANTS", "USE_FAST_ROUND", "NO_INLINE",
    "USE_LRU_CACHE", "LRU_CACHE_SIZE", "SHOULD_USE_CUDA",
]

# Module flags (RangeLibc.pyx:94-102 re-exports RangeLib.h:53-118).
USE_CACHED_TRIG = False
USE_ALTERNATE_MOD = True
USE_CACHED_CONSTANTS = True
USE_FAST_ROUND = False
NO_INLINE = False
USE_LRU_CACHE = False
LRU_CACHE_SIZE = 1000000
SHOULD_USE_CUDA = True  # the GPU class is always built here


def _png_occupancy(path: str, threshold: float) -> np.ndarray:
    """OMap(filename, threshold) (RangeLib.h:158-201) through the library's PNG
    decoder: x-major uint8 occupancy, 1 where the pixel's gray value is below
    the threshold."""
    import ctypes as C

    h = _lib.lib()
    w, hh = C.c_int(), C.c_int()
    _lib.check(h.rl_png_occupancy(path.encode(), threshold, None, C.byref(w), C.byref(hh)))
    occ = np.empty((w.value, hh.value), dtype=np.uint8)
    _lib.check(h.rl_png_occupancy(path.encode(), threshold, occ.ctypes.data, C.byref(w), C.byref(hh)))
    return occ


def ros_yaw(q) -> float:
    """Yaw of a quaternion message (x, y, z, w) as the reference computes it
    (RangeLibc.pyx:120-128): tf.transformations.euler_from_quaternion with the
    default static 'sxyz' axes, i.e. the quaternion's rotation matrix
    (normalised, scaled by sqrt(2 / |q|^2)) followed by
    atan2(M[1, 0], M[0, 0]), or 0 at the gimbal lock (|cos pitch| <= 4 eps)."""
    qv = np.array([float(q.x), float(q.y), float(q.z), float(q.w)], dtype=np.float64)
    eps4 = np.finfo(float).eps * 4.0
    nq = float(np.dot(qv, qv))
    if nq < eps4:
        return 0.0  # identity rotation
    qv *= math.sqrt(2.0 / nq)
    o = np.outer(qv, qv)
    m00 = 1.0 - o[1, 1] - o[2, 2]
    m10 = o[0, 1] + o[2, 3]
    cy = math.sqrt(m00 * m00 + m10 * m10)
    if cy > eps4:
        return math.atan2(m10, m00)
    return 0.0  # gimbal lock: the 'sxyz' decomposition puts the rotation in roll


class PyOMap:
    """Occupancy grid (RangeLibc.pyx:130-198 / OMap, RangeLib.h:121-322).

    PyOMap(width: int, height: int)     empty map
    PyOMap(path: str[, threshold])      PNG; occupied where gray < threshold (128)
    PyOMap(ndarray)                     ``grid[x][y] = arr[y, x] != 0``
    PyOMap(OccupancyGrid-like)          ROS map message (``.info``, ``.data``)

    The grid is held x-major: ``self.occ[x, y]`` with ``x < width``.  For a
    non-square ndarray the reference's size bookkeeping is undefined
    (RangeLibc.pyx:141-145 builds OMap(height, width)); here width is
    ``arr.shape[1]`` and height ``arr.shape[0]``.
    """

    def __init__(self, arg1, arg2=None):
        self._error = False
        self.fn = ""  # OMap::fn: the PNG path, empty otherwise (RangeLib.h:143, 159)
        self.world = (1.0, 0.0, 0.0, 0.0, 0.0, 1.0)  # RangeLibc.pyx:174-180
        if arg1 is not None and arg2 is not None:
            if isinstance(arg1, (int, np.integer)) and isinstance(arg2, (int, np.integer)):
                self.occ = np.zeros((int(arg1), int(arg2)), dtype=np.uint8)
            else:
                self._load_png(str(arg1), float(arg2))
        elif arg1 is not None:
            if isinstance(arg1, np.ndarray):
                if arg1.ndim != 2:
                    raise ValueError("PyOMap(ndarray) needs a 2-D array")
                self.occ = np.ascontiguousarray((np.asarray(arg1) != 0).T.astype(np.uint8))
            elif hasattr(arg1, "info") and hasattr(arg1, "data"):
                self._load_ros(arg1)
            else:
                self._load_png(str(arg1), 128.0)
        else:
            print("Failed to construct PyOMap, check argument types.")
            self.occ = np.zeros((1, 1), dtype=np.uint8)

    def _load_png(self, path: str, threshold: float):
        self.fn = path
        try:
            self.occ = _png_occupancy(path, threshold)
        except RangeLibError as e:  # OMap prints the decoder error and flags has_error
            print(f"ERROR: {e}")
            self._error = True
            self.occ = np.zeros((1, 1), dtype=np.uint8)

    def _load_ros(self, msg):
        """nav_msgs/OccupancyGrid (RangeLibc.pyx:146-167): cells > 10 occupied,
        grid[row][col]; world frame from the map origin."""
        info = msg.info
        width, height = int(info.width), int(info.height)
        arr = np.asarray(msg.data).reshape((height, width))
        self.occ = np.ascontiguousarray((arr > 10).astype(np.uint8))  # occ[x=row, y=col]
        angle = -1.0 * ros_yaw(info.origin.orientation)
        self.world = (float(info.resolution), angle, float(info.origin.position.x),
                      float(info.origin.position.y), float(np.sin(angle)), float(np.cos(angle)))

    def set_world(self, scale, angle, origin_x, origin_y, sin_angle=None, cos_angle=None):
        """Set the world-frame fields of OMap (RangeLib.h:134-141)."""
        s = math.sin(angle) if sin_angle is None else sin_angle
        c = math.cos(angle) if cos_angle is None else cos_angle
        self.world = (float(scale), float(angle), float(origin_x), float(origin_y), float(s), float(c))

    def save(self, fn) -> bool:
        """RGBA PNG, opaque, black where occupied and white elsewhere (OMap::save,
        RangeLib.h:264-291); returns False on success as the reference's error code."""
        occ = np.ascontiguousarray(self.occ, dtype=np.uint8)
        rc = _lib.lib().rl_png_save_occupancy(str(fn).encode(), occ.ctypes.data, occ.shape[0], occ.shape[1])
        if rc != 0:
            print(f"encoder error: {_lib.lib().rl_last_error().decode()}")
        return rc != 0

    def isOccupied(self, x: int, y: int) -> bool:
        return bool(self.occ[x, y])

    def error(self) -> bool:
        return self._error

    def width(self) -> int:
        return int(self.occ.shape[0])

    def height(self) -> int:
        return int(self.occ.shape[1])


def _need(a, dtype, ndim, name):
    """Typed-buffer check of the Cython signatures (np.ndarray[T, ndim=k, mode='c'])."""
    if not isinstance(a, np.ndarray):
        raise TypeError(f"{name}: expected numpy.ndarray")
    if a.dtype != dtype:
        raise ValueError(f"{name}: buffer dtype mismatch, expected {np.dtype(dtype)} got {a.dtype}")
    if a.ndim != ndim:
        raise ValueError(f"{name}: buffer has wrong number of dimensions (expected {ndim}, got {a.ndim})")
    if not a.flags.c_contiguous:
        raise ValueError(f"{name}: ndarray is not C-contiguous")
    if ndim == 2 and name == "ins" and a.shape[1] != 3:
        raise ValueError("ins: expected shape (n, 3)")
    if not a.flags.writeable and name in ("outs", "weights"):
        raise ValueError(f"{name}: buffer is read-only")


class _GpuMethod:
    METHOD = RL_METHOD_RM

    def __init__(self, Map: PyOMap, max_range: float, device: int = 0):
        if not isinstance(Map, PyOMap):
            raise TypeError("Map must be a PyOMap")
        self.max_range = float(np.float32(max_range))
        self.omap = Map
        self.map = DeviceMap(default_context(device), Map.occ)
        self.map.set_world(*Map.world)

    # RangeLibc.pyx:208-209, 325-326: calc_range_many converts world -> grid
    def calc_range_many(self, ins, outs):
        _need(ins, np.float32, 2, "ins")
        _need(outs, np.float32, 1, "outs")
        n = outs.shape[0]
        if ins.shape[0] < n:
            raise ValueError("ins has fewer rows than outs")
        self.map.numpy_calc_range(self.METHOD, self.max_range, ins, outs, n)

    def calc_range_many_grid(self, ins, outs):
        """RayMarchingGPU::calc_range_many of the C++ API: grid frame, no conversion."""
        _need(ins, np.float32, 2, "ins")
        _need(outs, np.float32, 1, "outs")
        n = outs.shape[0]
        if ins.shape[0] < n:
            raise ValueError("ins has fewer rows than outs")
        self.map.calc_range_many(self.METHOD, self.max_range, ins, outs, n)

    def calc_range_repeat_angles(self, ins, angles, outs):
        _need(ins, np.float32, 2, "ins")
        _need(angles, np.float32, 1, "angles")
        _need(outs, np.float32, 1, "outs")
        n, m = ins.shape[0], angles.shape[0]
        if outs.shape[0] < n * m:
            raise ValueError(f"outs needs {n * m} elements, has {outs.shape[0]}")
        self.map.numpy_calc_range_angles(self.METHOD, self.max_range, ins, angles, outs, n, m)

    def calc_range_repeat_angles_eval_sensor_model(self, ins, angles, obs, weights):
        _need(ins, np.float32, 2, "ins")
        _need(angles, np.float32, 1, "angles")
        _need(obs, np.float32, 1, "obs")
        _need(weights, np.float64, 1, "weights")
        n, m = ins.shape[0], angles.shape[0]
        if obs.shape[0] < m or weights.shape[0] < n:
            raise ValueError("obs needs one range per angle and weights one entry per particle")
        if not self.map.has_table:
            raise RangeLibError(3, "no sensor model set (call set_sensor_model first)")
        self.map.repeat_angles_eval_sensor_model(self.METHOD, self.max_range, ins, angles, obs, weights, n, m)

    def eval_sensor_model(self, observation, ranges, outs, num_rays, num_particles):
        _need(observation, np.float32, 1, "observation")
        _need(ranges, np.float32, 1, "ranges")
        _need(outs, np.float64, 1, "outs")
        if observation.shape[0] < num_rays or ranges.shape[0] < num_rays * num_particles \
                or outs.shape[0] < num_particles:
            raise ValueError("array sizes do not cover num_rays x num_particles")
        if not self.map.has_table:
            raise RangeLibError(3, "no sensor model set (call set_sensor_model first)")
        self.map.eval_sensor_model(observation, ranges, outs, int(num_rays), int(num_particles))

    def set_sensor_model(self, table):
        _need(table, np.float64, 2, "table")
        if not table.shape[0] == table.shape[1]:
            print("Sensor model must have equal matrix dimensions, failing!")
            return
        self.map.set_sensor_table(table)

    def saveTrace(self, path):
        print("WARNING: trace map not generated, must compile with trace support enabled.")


class PyRayMarchingGPU(_GpuMethod):
    """RayMarchingGPU (RangeLib.h:774-920; RangeLibc.pyx:314-340)."""

    METHOD = RL_METHOD_RM


class PyRayMarching(_GpuMethod):
    """RayMarching (RangeLib.h:922-969; RangeLibc.pyx:227-251), run on the GPU."""

    METHOD = RL_METHOD_RM

    def calc_range(self, x, y, heading):
        return self.map.calc_range(self.METHOD, self.max_range, float(x), float(y), float(heading))


class PyBresenhamsLine(_GpuMethod):
    """BresenhamsLine (RangeLib.h:691-772; RangeLibc.pyx:200-225), run on the GPU."""

    METHOD = RL_METHOD_BL

    def calc_range(self, x, y, heading):
        return self.map.calc_range(self.METHOD, self.max_range, float(x), float(y), float(heading))


def _g7(v) -> str:
    """A float as std::ostream prints it at setprecision(7): C's %.7g of the value."""
    return "%.7g" % float(np.float32(v))


def _vec(vals) -> str:
    """utils::serialize(std::vector<float>) (RangeUtils.h:222-232)."""
    return "[" + ",".join(_g7(v) for v in vals) + "]"


def cddt_json(theta_disc, trans, max_range, fn, occ_xmajor, widths, start, vals) -> str:
    """CDDTCast::serializeJson (RangeLib.h:1690-1735) from CSR tables: the same text,
    line by line (J1 = two spaces; map data x-major as grid[x]; theta i * (2 pi / td))."""
    td = int(theta_disc)
    occ = np.asarray(occ_xmajor)
    w, h = occ.shape
    step = np.float32(6.28318530718 / np.float64(np.float32(td)))
    out = ['{"cddt": {\n', '  "theta_discretization": %d,\n' % td, '  "lut_translations": ' + _vec(trans) + ",\n",
           '  "max_range":' + _g7(max_range) + ",\n", '  "map": {\n', '    "path": "' + fn + '",\n',
           '    "width": %d,\n' % w, '    "height": %d,\n' % h,
           '    "data": [' + ",".join("[" + ",".join("1" if c else "0" for c in occ[x]) + "]" for x in range(w)) + "],\n",
           "  },\n", '  "compressed_lut": [\n']
    b = 0
    for i in range(td):
        out.append("    {\n")
        out.append('      "theta": ' + _g7(np.float32(i) * step) + ",\n")
        bins = []
        for _ in range(int(widths[i])):
            bins.append(_vec(vals[start[b]:start[b + 1]]))
            b += 1
        out.append('      "zeros": [' + ",".join(bins) + "]\n")
        out.append("    },\n" if i < td - 1 else "    }\n")
    out.append("  ]\n")
    out.append("}}\n")
    return "".join(out)


class PyCDDTCast(_GpuMethod):
    """CDDTCast (RangeLib.h:971-1770; RangeLibc.pyx:253-287): compressed directional
    distance transform, tables built on the GPU at construction; prune() makes it PCDDT."""

    METHOD = RL_METHOD_CDDT

    def __init__(self, Map: PyOMap, max_range: float, theta_disc: int, device: int = 0):
        super().__init__(Map, max_range, device)
        self.theta_disc = int(theta_disc)
        self.map.build_cddt(self.theta_disc)

    def prune(self, max_range: float = -1.0):
        """RangeLibc.pyx:261-265: a negative max_range prunes at the constructor's."""
        self.map.prune_cddt(self.max_range if max_range < 0.0 else float(np.float32(max_range)))

    def serializeJson(self) -> str:
        """CDDTCast::serializeJson (RangeLib.h:1690-1735), the text the reference writes."""
        t = self.map.cddt_tables()
        return cddt_json(self.theta_disc, t["trans"], self.max_range, self.omap.fn, self.omap.occ, t["widths"],
                         t["start"], t["vals"])

    def calc_range_many_radial_optimized(self, num_rays, min_angle, max_angle, ins, outs):
        """RangeLibc.pyx:273-275: num_rays beams over [min_angle, max_angle] per pose,
        opposite beams answered by one pair query."""
        _need(ins, np.float32, 2, "ins")
        _need(outs, np.float32, 1, "outs")
        n = ins.shape[0]
        if outs.shape[0] < n * num_rays:
            raise ValueError(f"outs needs {n * num_rays} elements, has {outs.shape[0]}")
        self.map.calc_range_many_radial_optimized(self.METHOD, self.max_range, ins, outs, n, int(num_rays),
                                                  float(min_angle), float(max_angle))

    def calc_range(self, x, y, heading):
        return self.map.calc_range(self.METHOD, self.max_range, float(x), float(y), float(heading))


class PyGiantLUTCast(_GpuMethod):
    """GiantLUTCast (RangeLib.h:1772-1900; RangeLibc.pyx:289-312): the table is built on
    the GPU at construction, one marched ray per (x, y, theta bin)."""

    METHOD = RL_METHOD_GLT

    def __init__(self, Map: PyOMap, max_range: float, theta_disc: int, device: int = 0):
        super().__init__(Map, max_range, device)
        self.theta_disc = int(theta_disc)
        self.map.build_giant_lut(self.max_range, self.theta_disc)

    def calc_range(self, x, y, heading):
        return self.map.calc_range(self.METHOD, self.max_range, float(x), float(y), float(heading))

    def lut_size(self):
        """Bytes of the table (RangeLib.h:1818-1820)."""
        return self.map.width * self.map.height * self.theta_disc * 2


class PyNull:
    """PyNull (RangeLibc.pyx:342-351): a no-op method for call-overhead timing."""

    def __init__(self, Map, max_range, theta_disc=0):
        pass

    def calc_range(self, x, y, heading):
        return x + y + heading

    def calc_range_many(self, ins, outs):
        return None
