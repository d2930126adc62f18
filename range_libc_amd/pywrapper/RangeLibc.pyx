# cython: language_level=3, boundscheck=False, wraparound=False
# distutils: language = c
"""range_libc -- the Cython module of the reference (pywrapper/RangeLibc.pyx),
bound to the C ABI of rangelib_hip.h instead of RangeLib.h.

Same module name, class names, constructor arguments, cpdef method names,
argument types (typed C-contiguous ndarrays) and argument order as
RangeLibc.pyx:130-351; every casting method runs on the GPU and releases the
GIL around its call.  Differences: errors raise RuntimeError carrying
rl_last_error() (the reference prints or aborts); ROS messages are accepted
duck-typed (.info, .data) without importing ROS; strings are str.

    import sys; sys.path.insert(0, "<repo>/range_libc_amd/pywrapper")
    import range_libc
    omap = range_libc.PyOMap("map.png", 128)      # or an ndarray / OccupancyGrid
    rm = range_libc.PyRayMarchingGPU(omap, 500)
    rm.calc_range_repeat_angles(poses, angles, ranges)
"""
import math

import numpy as np
cimport numpy as np
from libc.stdint cimport uint8_t

np.import_array()

cdef extern from "rangelib_hip.h":
    ctypedef struct rl_ctx:
        pass
    ctypedef struct rl_map:
        pass
    int RL_METHOD_RM
    int RL_METHOD_BL
    int RL_METHOD_GLT
    int RL_METHOD_CDDT
    const char* rl_last_error()
    int rl_create(int device, rl_ctx** out)
    int rl_map_create(rl_ctx* ctx, const uint8_t* occ, int width, int height, rl_map** out)
    int rl_map_destroy(rl_map* m)
    int rl_map_set_world(rl_map* m, float scale, float angle, float ox, float oy, float s, float c)
    int rl_map_set_sensor_table(rl_map* m, const double* table, int k)
    int rl_map_build_cddt(rl_map* m, int theta_disc)
    int rl_map_prune_cddt(rl_map* m, float max_range)
    int rl_map_build_giant_lut(rl_map* m, float max_range, int theta_disc)
    int rl_calc_range(rl_map* m, int method, float max_range, float x, float y, float heading, float* out) nogil
    int rl_numpy_calc_range(rl_map* m, int method, float max_range, const float* ins, float* outs, int n) nogil
    int rl_numpy_calc_range_angles(rl_map* m, int method, float max_range, const float* ins,
                                   const float* angles, float* outs, int num_particles, int num_angles) nogil
    int rl_eval_sensor_model(rl_map* m, const float* obs, const float* ranges, double* outs,
                             int rays_per_particle, int particles) nogil
    int rl_calc_range_repeat_angles_eval_sensor_model(rl_map* m, int method, float max_range, const float* ins,
                                                      const float* angles, const float* obs, double* weights,
                                                      int num_particles, int num_angles) nogil
    int rl_calc_range_many_radial_optimized(rl_map* m, int method, float max_range, const float* ins, float* outs,
                                            int num_particles, int num_rays, float min_angle,
                                            float max_angle) nogil
    int rl_png_occupancy(const char* path, float threshold, uint8_t* occ_xmajor, int* width, int* height)
    int rl_png_save_occupancy(const char* path, const uint8_t* occ_xmajor, int width, int height)

# module flags (RangeLibc.pyx:94-102 with RangeLib.h:53-118's defaults)
USE_CACHED_TRIG = False
USE_ALTERNATE_MOD = True
USE_CACHED_CONSTANTS = True
USE_FAST_ROUND = False
NO_INLINE = False
USE_LRU_CACHE = False
LRU_CACHE_SIZE = 1000000
SHOULD_USE_CUDA = True

cdef rl_ctx* _ctx = NULL


cdef int _check(int rc) except -1:
    if rc != 0:
        raise RuntimeError("rangelib_hip: " + rl_last_error().decode())
    return 0


cdef rl_ctx* _context() except NULL:
    global _ctx
    if _ctx == NULL:
        _check(rl_create(0, &_ctx))
    return _ctx


def quaternion_to_angle(q):
    """Yaw of a quaternion message (RangeLibc.pyx:120-128): the reference calls
    tf.transformations.euler_from_quaternion (static 'sxyz' axes), i.e.
    atan2(M[1, 0], M[0, 0]) of the normalised quaternion's rotation matrix, or
    0 at the gimbal lock (|cos pitch| <= 4 eps)."""
    qv = np.array([float(q.x), float(q.y), float(q.z), float(q.w)], dtype=np.float64)
    eps4 = np.finfo(float).eps * 4.0
    nq = float(np.dot(qv, qv))
    if nq < eps4:
        return 0.0
    qv *= math.sqrt(2.0 / nq)
    o = np.outer(qv, qv)
    m00 = 1.0 - o[1, 1] - o[2, 2]
    m10 = o[0, 1] + o[2, 3]
    cy = math.sqrt(m00 * m00 + m10 * m10)
    if cy > eps4:
        return math.atan2(m10, m00)
    return 0.0


cdef class PyOMap:
    """OMap (RangeLib.h:121-322) as RangeLibc.pyx:130-198 builds it:
    PyOMap(width, height), PyOMap(path[, threshold]), PyOMap(ndarray),
    PyOMap(OccupancyGrid).  The grid is held x-major (occ[x, y])."""
    cdef public object occ
    cdef public str fn
    cdef bint has_error
    cdef public float world_scale, world_angle, world_origin_x, world_origin_y, world_sin_angle, world_cos_angle

    def __cinit__(self, arg1, arg2=None):
        set_trans_params = False
        self.fn = ""
        self.has_error = False
        if arg1 is not None and arg2 is not None:
            if isinstance(arg1, (int, np.integer)) and isinstance(arg2, (int, np.integer)):
                self.occ = np.zeros((int(arg1), int(arg2)), dtype=np.uint8)
            else:
                self._load_png(str(arg1), float(arg2))
        elif arg1 is not None:
            if isinstance(arg1, np.ndarray):
                # grid[x][y] = arr[y, x] (RangeLibc.pyx:141-145); width = shape[1]
                self.occ = np.ascontiguousarray((np.asarray(arg1) != 0).T.astype(np.uint8))
            elif hasattr(arg1, "info") and hasattr(arg1, "data"):
                info = arg1.info
                width, height = int(info.width), int(info.height)
                array_255 = np.array(arg1.data).reshape((height, width))
                # OMap(height, width) with grid[x][y] = data[x, y] > 10 (RangeLibc.pyx:148-157)
                self.occ = np.ascontiguousarray((array_255 > 10).astype(np.uint8))
                angle = -1.0 * quaternion_to_angle(info.origin.orientation)
                self.world_scale = info.resolution
                self.world_angle = angle
                self.world_origin_x = info.origin.position.x
                self.world_origin_y = info.origin.position.y
                self.world_sin_angle = np.sin(angle)
                self.world_cos_angle = np.cos(angle)
                set_trans_params = True
            else:
                self._load_png(str(arg1), 128.0)
        else:
            print("Failed to construct PyOMap, check argument types.")
            self.occ = np.zeros((1, 1), dtype=np.uint8)
        if not set_trans_params:
            self.world_scale = 1.0
            self.world_angle = 0.0
            self.world_origin_x = 0.0
            self.world_origin_y = 0.0
            self.world_sin_angle = 0.0
            self.world_cos_angle = 1.0

    def _load_png(self, str path, float threshold):
        cdef int w = 0, h = 0
        cdef np.ndarray[np.uint8_t, ndim=2, mode="c"] occ
        self.fn = path
        bpath = path.encode()
        rc = rl_png_occupancy(bpath, threshold, NULL, &w, &h)
        if rc == 0:
            occ = np.empty((w, h), dtype=np.uint8)
            rc = rl_png_occupancy(bpath, threshold, &occ[0, 0], &w, &h)
        if rc != 0:  # OMap prints the decoder error and flags has_error (RangeLib.h:164-168)
            print("ERROR %d: %s" % (rc, rl_last_error().decode()))
            self.has_error = True
            self.occ = np.zeros((1, 1), dtype=np.uint8)
            return
        self.occ = occ

    cpdef bint save(self, str fn):
        cdef np.ndarray[np.uint8_t, ndim=2, mode="c"] occ = np.ascontiguousarray(self.occ, dtype=np.uint8)
        rc = rl_png_save_occupancy(fn.encode(), &occ[0, 0], occ.shape[0], occ.shape[1])
        if rc != 0:
            print("encoder error %d: %s" % (rc, rl_last_error().decode()))
        return rc != 0

    cpdef bint isOccupied(self, int x, int y):
        return self.occ[x, y] != 0

    cpdef bint error(self):
        return self.has_error

    cpdef int width(self):
        return self.occ.shape[0]

    cpdef int height(self):
        return self.occ.shape[1]


cdef class _Method:
    cdef rl_map* m
    cdef float max_range
    cdef int method

    cdef int _open(self, PyOMap Map, float max_range, int method) except -1:
        cdef np.ndarray[np.uint8_t, ndim=2, mode="c"] occ = np.ascontiguousarray(Map.occ, dtype=np.uint8)
        self.max_range = max_range
        self.method = method
        _check(rl_map_create(_context(), &occ[0, 0], occ.shape[0], occ.shape[1], &self.m))
        _check(rl_map_set_world(self.m, Map.world_scale, Map.world_angle, Map.world_origin_x, Map.world_origin_y,
                                Map.world_sin_angle, Map.world_cos_angle))
        return 0

    def __dealloc__(self):
        if self.m != NULL:
            rl_map_destroy(self.m)
            self.m = NULL

    cdef float _single(self, float x, float y, float heading) except? -1:
        cdef float out = 0
        _check(rl_calc_range(self.m, self.method, self.max_range, x, y, heading, &out))
        return out

    cpdef void calc_range_many(self, np.ndarray[float, ndim=2, mode="c"] ins, np.ndarray[float, ndim=1, mode="c"] outs):
        cdef int n = outs.shape[0], rc
        if ins.shape[0] < n:
            raise ValueError("ins has fewer rows than outs")
        with nogil:
            rc = rl_numpy_calc_range(self.m, self.method, self.max_range, &ins[0, 0], &outs[0], n)
        _check(rc)

    cpdef void calc_range_repeat_angles(self, np.ndarray[float, ndim=2, mode="c"] ins,
                                        np.ndarray[float, ndim=1, mode="c"] angles,
                                        np.ndarray[float, ndim=1, mode="c"] outs):
        cdef int n = ins.shape[0], na = angles.shape[0], rc
        if outs.shape[0] < n * na:
            raise ValueError("outs needs num_particles * num_angles elements")
        with nogil:
            rc = rl_numpy_calc_range_angles(self.m, self.method, self.max_range, &ins[0, 0], &angles[0], &outs[0], n,
                                            na)
        _check(rc)

    cpdef void calc_range_repeat_angles_eval_sensor_model(self, np.ndarray[float, ndim=2, mode="c"] ins,
                                                          np.ndarray[float, ndim=1, mode="c"] angles,
                                                          np.ndarray[float, ndim=1, mode="c"] obs,
                                                          np.ndarray[double, ndim=1, mode="c"] weights):
        cdef int n = ins.shape[0], na = angles.shape[0], rc
        if obs.shape[0] < na or weights.shape[0] < n:
            raise ValueError("obs needs one range per angle and weights one entry per particle")
        with nogil:
            rc = rl_calc_range_repeat_angles_eval_sensor_model(self.m, self.method, self.max_range, &ins[0, 0],
                                                               &angles[0], &obs[0], &weights[0], n, na)
        _check(rc)

    cpdef void eval_sensor_model(self, np.ndarray[float, ndim=1, mode="c"] observation,
                                 np.ndarray[float, ndim=1, mode="c"] ranges,
                                 np.ndarray[double, ndim=1, mode="c"] outs, int num_rays, int num_particles):
        cdef int rc
        if observation.shape[0] < num_rays or ranges.shape[0] < num_rays * num_particles \
                or outs.shape[0] < num_particles:
            raise ValueError("array sizes do not cover num_rays x num_particles")
        with nogil:
            rc = rl_eval_sensor_model(self.m, &observation[0], &ranges[0], &outs[0], num_rays, num_particles)
        _check(rc)

    cpdef void set_sensor_model(self, np.ndarray[double, ndim=2, mode="c"] table):
        if not table.shape[0] == table.shape[1]:
            print("Sensor model must have equal matrix dimensions, failing!")
            return
        _check(rl_map_set_sensor_table(self.m, &table[0, 0], table.shape[0]))


cdef class PyBresenhamsLine(_Method):
    """BresenhamsLine (RangeLib.h:691-772; RangeLibc.pyx:200-225)."""
    def __cinit__(self, PyOMap Map, float max_range):
        self._open(Map, max_range, RL_METHOD_BL)

    cpdef float calc_range(self, float x, float y, float heading):
        return self._single(x, y, heading)

    cpdef float saveTrace(self, str path):
        print("WARNING: trace map not generated, must compile with trace support enabled.")


cdef class PyRayMarching(_Method):
    """RayMarching (RangeLib.h:922-969; RangeLibc.pyx:227-251)."""
    def __cinit__(self, PyOMap Map, float max_range):
        self._open(Map, max_range, RL_METHOD_RM)

    cpdef float calc_range(self, float x, float y, float heading):
        return self._single(x, y, heading)

    cpdef float saveTrace(self, str path):
        print("WARNING: trace map not generated, must compile with trace support enabled.")


cdef class PyCDDTCast(_Method):
    """CDDTCast (RangeLib.h:971-1770; RangeLibc.pyx:253-287): tables built on the GPU."""
    def __cinit__(self, PyOMap Map, float max_range, unsigned int theta_disc):
        self._open(Map, max_range, RL_METHOD_CDDT)
        _check(rl_map_build_cddt(self.m, theta_disc))

    cpdef void prune(self, float max_range=-1.0):
        _check(rl_map_prune_cddt(self.m, self.max_range if max_range < 0.0 else max_range))

    cpdef float calc_range(self, float x, float y, float heading):
        return self._single(x, y, heading)

    cpdef void calc_range_many_radial_optimized(self, int num_rays, float min_angle, float max_angle,
                                                np.ndarray[float, ndim=2, mode="c"] ins,
                                                np.ndarray[float, ndim=1, mode="c"] outs):
        cdef int n = ins.shape[0], rc
        if outs.shape[0] < n * num_rays:
            raise ValueError("outs needs num_particles * num_rays elements")
        with nogil:
            rc = rl_calc_range_many_radial_optimized(self.m, self.method, self.max_range, &ins[0, 0], &outs[0], n,
                                                     num_rays, min_angle, max_angle)
        _check(rc)


cdef class PyGiantLUTCast(_Method):
    """GiantLUTCast (RangeLib.h:1772-1900; RangeLibc.pyx:289-312): table built on the GPU."""
    def __cinit__(self, PyOMap Map, float max_range, unsigned int theta_disc):
        self._open(Map, max_range, RL_METHOD_GLT)
        _check(rl_map_build_giant_lut(self.m, max_range, theta_disc))

    cpdef float calc_range(self, float x, float y, float heading):
        return self._single(x, y, heading)


cdef class PyRayMarchingGPU(_Method):
    """RayMarchingGPU (RangeLib.h:774-920; RangeLibc.pyx:314-340)."""
    def __cinit__(self, PyOMap Map, float max_range):
        self._open(Map, max_range, RL_METHOD_RM)


cdef class PyNull:
    """PyNull (RangeLibc.pyx:342-351): a stand-in that casts nothing."""
    def __cinit__(self, PyOMap Map, float max_range, unsigned int theta_disc):
        pass

    cpdef float calc_range(self, float x, float y, float heading):
        return x + y + heading

    cpdef void calc_range_many(self, np.ndarray[float, ndim=2, mode="c"] ins, np.ndarray[float, ndim=1, mode="c"] outs):
        a = ins[0, 0]
        b = outs[0]
