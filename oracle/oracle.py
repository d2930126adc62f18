"""ctypes front end of the CPU oracle -- TEST INFRASTRUCTURE ONLY.

``Oracle`` wraps ``liboracle.so`` (the C restatement in rl_oracle.c) and
``Reference`` wraps ``_ref/libref.so`` (the reference's own RangeLib.h compiled
from its source tree, see Makefile).  Only the tests, the smoke check and the
benchmark's CPU-baseline leg use this module.

Arrays follow the reference's conventions: occupancy and distance transform
are x-major (``a[x*height + y]``, i.e. the transpose of an image ``img[y, x]``),
poses are C-contiguous float32 (n, 3), tables float64 row-major.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libref.so")

_fp = C.POINTER(C.c_float)
_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_u8p = C.POINTER(C.c_uint8)


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _ptr(a, t):
    return a.ctypes.data_as(t)


def xmajor(img: np.ndarray) -> np.ndarray:
    """Image ``img[y, x]`` -> x-major uint8 occupancy ``occ[x, y]``."""
    return np.ascontiguousarray(np.asarray(img, dtype=np.uint8).T)


class _World(C.Structure):
    _fields_ = [(n, C.c_float) for n in ("scale", "angle", "origin_x", "origin_y", "sin_angle", "cos_angle")]


_llp = C.POINTER(C.c_longlong)


class _Cddt(C.Structure):
    _fields_ = [("td", C.c_int), ("nang", C.c_int), ("width", C.POINTER(C.c_int)), ("trans", _fp), ("cosv", _fp),
                ("sinv", _fp), ("bin_off", _llp), ("start", _llp), ("vals", _fp), ("td_div_2pi", C.c_float),
                ("m2pi_div_td", C.c_float)]


class _Method(C.Structure):
    _fields_ = [("kind", C.c_int), ("width", C.c_int), ("height", C.c_int), ("occ", _u8p),
                ("edt", _fp), ("max_range", C.c_float), ("world", _World), ("table", _dp),
                ("table_width", C.c_int), ("glt", C.POINTER(C.c_uint16)), ("glt_td", C.c_int),
                ("glt_max_div_limits", C.c_float), ("glt_td_div_2pi", C.c_float), ("cddt", C.POINTER(_Cddt))]


IDENTITY_WORLD = (1.0, 0.0, 0.0, 0.0, 0.0, 1.0)


def _load_oracle():
    if not os.path.exists(ORACLE_SO):
        raise RuntimeError(f"oracle library missing: {ORACLE_SO} (run `make -C oracle`)")
    lib = C.CDLL(ORACLE_SO)
    lib.or_edt.argtypes = [_u8p, C.c_int, C.c_int, _fp]
    lib.or_rm_calc_range.argtypes = [_fp, C.c_int, C.c_int, C.c_float, C.c_float, C.c_float, C.c_float, _ip]
    lib.or_rm_calc_range.restype = C.c_float
    lib.or_bl_calc_range.argtypes = [_u8p, C.c_int, C.c_int, C.c_float, C.c_float, C.c_float, C.c_float, _ip]
    lib.or_bl_calc_range.restype = C.c_float
    mp = C.POINTER(_Method)
    lib.or_calc_range.argtypes = [mp, C.c_float, C.c_float, C.c_float]
    lib.or_calc_range.restype = C.c_float
    lib.or_calc_range_many.argtypes = [mp, _fp, _fp, C.c_int, _ip]
    lib.or_numpy_calc_range.argtypes = [mp, _fp, _fp, C.c_int, _ip]
    lib.or_numpy_calc_range_angles.argtypes = [mp, _fp, _fp, _fp, C.c_int, C.c_int, _ip]
    lib.or_eval_sensor_model.argtypes = [mp, _fp, _fp, _dp, C.c_int, C.c_int]
    lib.or_calc_range_repeat_angles_eval_sensor_model.argtypes = [mp, _fp, _fp, _fp, _dp, C.c_int, C.c_int]
    lib.or_glt_build.argtypes = [_fp, C.c_int, C.c_int, C.c_float, C.c_int, C.POINTER(C.c_uint16), _fp]
    lib.or_glt_discretize.argtypes = [C.c_float, C.c_int, C.c_float]
    lib.or_glt_discretize.restype = C.c_int
    lib.or_glt_calc_range.argtypes = [mp, C.c_float, C.c_float, C.c_float]
    lib.or_glt_calc_range.restype = C.c_float
    lib.or_cddt_build.argtypes = [_u8p, C.c_int, C.c_int, C.c_int]
    lib.or_cddt_build.restype = C.POINTER(_Cddt)
    lib.or_cddt_free.argtypes = [C.POINTER(_Cddt)]
    lib.or_cddt_prune.argtypes = [C.POINTER(_Cddt), _u8p, C.c_int, C.c_int, C.c_float]
    lib.or_calc_range_many_radial_optimized.argtypes = [mp, _fp, _fp, C.c_int, C.c_int, C.c_float, C.c_float]
    lib.or_sinf_many.argtypes = [_fp, _fp, C.c_long]
    lib.or_cosf_many.argtypes = [_fp, _fp, C.c_long]
    return lib


_ORACLE = None


def oracle_lib():
    global _ORACLE
    if _ORACLE is None:
        _ORACLE = _load_oracle()
    return _ORACLE


def edt(img: np.ndarray) -> np.ndarray:
    """Distance transform of image ``img[y, x]``; returned x-major (W, H) float32."""
    occ = xmajor(img)
    w, h = occ.shape
    out = np.empty((w, h), dtype=np.float32)
    oracle_lib().or_edt(_ptr(occ, _u8p), w, h, _ptr(out, _fp))
    return out


def libc_sinf(x):
    x = _f32(x)
    out = np.empty_like(x)
    oracle_lib().or_sinf_many(_ptr(x, _fp), _ptr(out, _fp), x.size)
    return out


def libc_cosf(x):
    x = _f32(x)
    out = np.empty_like(x)
    oracle_lib().or_cosf_many(_ptr(x, _fp), _ptr(out, _fp), x.size)
    return out


def glt_build(edt_xmajor: np.ndarray, max_range: float, theta_disc: int):
    """GiantLUTCast table (W, H, td) uint16 and (max_div_limits, limits_div_max, td_div_2pi)."""
    e = np.ascontiguousarray(edt_xmajor, dtype=np.float32)
    w, h = e.shape
    table = np.empty((w, h, theta_disc), dtype=np.uint16)
    consts = np.empty(3, dtype=np.float32)
    oracle_lib().or_glt_build(_ptr(e, _fp), w, h, max_range, theta_disc,
                              table.ctypes.data_as(C.POINTER(C.c_uint16)), _ptr(consts, _fp))
    return table, consts


def glt_discretize(theta: float, theta_disc: int, td_div_2pi: float) -> int:
    return oracle_lib().or_glt_discretize(theta, theta_disc, td_div_2pi)


def cddt_tables(c) -> dict:
    """CSR view (copies) of an or_cddt: widths, translations, start, vals."""
    cd = c.contents
    td = cd.td
    widths = np.ctypeslib.as_array(cd.width, (td,)).copy()
    nb = int(widths.sum())
    start = np.ctypeslib.as_array(cd.start, (nb + 1,)).copy()
    vals = np.ctypeslib.as_array(cd.vals, (max(int(start[-1]), 1),))[: int(start[-1])].copy()
    return {"widths": widths, "trans": np.ctypeslib.as_array(cd.trans, (td,)).copy(), "start": start, "vals": vals}


class Oracle:
    """One range method (``kind`` 'rm', 'bl' or 'glt') bound to one map, as
    ``RayMarching(OMap, max_range)`` / ``BresenhamsLine(OMap, max_range)`` /
    ``GiantLUTCast(OMap, max_range, theta_disc)``."""

    def __init__(self, img: np.ndarray, max_range: float, kind: str = "rm", world=IDENTITY_WORLD,
                 theta_disc: int = 0, prune: bool = False):
        self.lib = oracle_lib()
        self.occ = xmajor(img)
        self.width, self.height = self.occ.shape
        self.edt = edt(img) if kind in ("rm", "glt") else np.zeros((1, 1), dtype=np.float32)
        self.table = np.zeros((1, 1), dtype=np.float64)
        self.m = _Method()
        self.m.kind = {"rm": 0, "bl": 1, "glt": 2, "cddt": 3}[kind]
        self._cddt = None
        if kind == "cddt":
            self._cddt = self.lib.or_cddt_build(_ptr(self.occ, _u8p), self.width, self.height, theta_disc)
            if prune:
                self.lib.or_cddt_prune(self._cddt, _ptr(self.occ, _u8p), self.width, self.height, max_range)
            self.m.cddt = self._cddt
        if kind == "glt":
            self.glt, consts = glt_build(self.edt, max_range, theta_disc)
            self.m.glt = self.glt.ctypes.data_as(C.POINTER(C.c_uint16))
            self.m.glt_td = theta_disc
            self.m.glt_max_div_limits = float(consts[0])
            self.m.glt_td_div_2pi = float(consts[2])
        self.m.width = self.width
        self.m.height = self.height
        self.m.occ = _ptr(self.occ, _u8p)
        self.m.edt = _ptr(self.edt, _fp)
        self.m.max_range = max_range
        self.m.world = _World(*[float(v) for v in world])
        self.m.table = _ptr(self.table, _dp)
        self.m.table_width = 1

    def __del__(self):
        if getattr(self, "_cddt", None):
            self.lib.or_cddt_free(self._cddt)
            self._cddt = None

    def cddt_tables(self) -> dict:
        return cddt_tables(self._cddt)

    def calc_range_many_radial_optimized(self, ins, num_rays, min_angle, max_angle, fill=-7.0):
        ins = _f32(ins)
        outs = np.full(ins.shape[0] * num_rays, fill, dtype=np.float32)
        self.lib.or_calc_range_many_radial_optimized(C.byref(self.m), _ptr(ins, _fp), _ptr(outs, _fp), ins.shape[0],
                                                     num_rays, min_angle, max_angle)
        return outs

    def set_sensor_model(self, table: np.ndarray):
        self.table = np.ascontiguousarray(table, dtype=np.float64)
        assert self.table.shape[0] == self.table.shape[1]
        self.m.table = _ptr(self.table, _dp)
        self.m.table_width = self.table.shape[0]

    def calc_range(self, x, y, heading):
        return self.lib.or_calc_range(C.byref(self.m), x, y, heading)

    def calc_range_many(self, ins, counts=False):
        """Grid-frame batch (no world conversion)."""
        ins = _f32(ins)
        n = ins.shape[0]
        outs = np.empty(n, dtype=np.float32)
        cnt = np.zeros(n, dtype=np.int32)
        self.lib.or_calc_range_many(C.byref(self.m), _ptr(ins, _fp), _ptr(outs, _fp), n, _ptr(cnt, _ip))
        return (outs, cnt) if counts else outs

    def numpy_calc_range(self, ins, counts=False):
        ins = _f32(ins)
        n = ins.shape[0]
        outs = np.empty(n, dtype=np.float32)
        cnt = np.zeros(n, dtype=np.int32)
        self.lib.or_numpy_calc_range(C.byref(self.m), _ptr(ins, _fp), _ptr(outs, _fp), n, _ptr(cnt, _ip))
        return (outs, cnt) if counts else outs

    def numpy_calc_range_angles(self, ins, angles, counts=False):
        ins, angles = _f32(ins), _f32(angles)
        n, m = ins.shape[0], angles.shape[0]
        outs = np.empty(n * m, dtype=np.float32)
        cnt = np.zeros(n * m, dtype=np.int32)
        self.lib.or_numpy_calc_range_angles(C.byref(self.m), _ptr(ins, _fp), _ptr(angles, _fp),
                                            _ptr(outs, _fp), n, m, _ptr(cnt, _ip))
        return (outs, cnt) if counts else outs

    def eval_sensor_model(self, obs, ranges, num_rays, num_particles):
        obs, ranges = _f32(obs), _f32(ranges)
        outs = np.empty(num_particles, dtype=np.float64)
        self.lib.or_eval_sensor_model(C.byref(self.m), _ptr(obs, _fp), _ptr(ranges, _fp), _ptr(outs, _dp),
                                      num_rays, num_particles)
        return outs

    def calc_range_repeat_angles_eval_sensor_model(self, ins, angles, obs):
        ins, angles, obs = _f32(ins), _f32(angles), _f32(obs)
        n, m = ins.shape[0], angles.shape[0]
        w = np.empty(n, dtype=np.float64)
        self.lib.or_calc_range_repeat_angles_eval_sensor_model(
            C.byref(self.m), _ptr(ins, _fp), _ptr(angles, _fp), _ptr(obs, _fp), _ptr(w, _dp), n, m)
        return w


# ---------------------------------------------------------------------------
# The compiled reference (optional: present only where it was built).

def ref_available() -> bool:
    return os.path.exists(REF_SO)


_REF = None


def ref_lib():
    global _REF
    if _REF is None:
        if not ref_available():
            raise RuntimeError(f"compiled reference missing: {REF_SO} (run `make -C oracle ref`)")
        lib = C.CDLL(REF_SO)
        lib.ref_load_png.argtypes = [C.c_char_p, C.c_float, _ip, _ip, _u8p]
        lib.ref_load_png.restype = C.c_int
        lib.ref_edt.argtypes = [_u8p, C.c_int, C.c_int, _fp]
        lib.ref_create.argtypes = [C.c_int, _u8p, C.c_int, C.c_int, C.c_float, _fp]
        lib.ref_create.restype = C.c_void_p
        lib.ref_create_glt.argtypes = [_u8p, C.c_int, C.c_int, C.c_float, C.c_int, _fp]
        lib.ref_create_glt.restype = C.c_void_p
        lib.ref_glt_lut.argtypes = [C.c_void_p, C.POINTER(C.c_uint16), _fp]
        lib.ref_create_cddt.argtypes = [_u8p, C.c_int, C.c_int, C.c_float, C.c_int, _fp]
        lib.ref_create_cddt.restype = C.c_void_p
        lib.ref_cddt_prune.argtypes = [C.c_void_p, C.c_float]
        lib.ref_cddt_sizes.argtypes = [C.c_void_p, _llp, _llp]
        lib.ref_cddt_export.argtypes = [C.c_void_p, C.POINTER(C.c_int), _fp, _llp, _fp]
        lib.ref_calc_range_many_radial_optimized.argtypes = [C.c_void_p, _fp, _fp, C.c_int, C.c_int, C.c_float,
                                                             C.c_float]
        lib.ref_cddt_serialize_json.argtypes = [C.c_void_p, C.c_char_p, C.c_longlong]
        lib.ref_cddt_serialize_json.restype = C.c_longlong
        lib.ref_destroy.argtypes = [C.c_void_p]
        lib.ref_set_sensor_model.argtypes = [C.c_void_p, _dp, C.c_int]
        lib.ref_calc_range.argtypes = [C.c_void_p, C.c_float, C.c_float, C.c_float]
        lib.ref_calc_range.restype = C.c_float
        lib.ref_calc_range_many.argtypes = [C.c_void_p, _fp, _fp, C.c_int]
        lib.ref_numpy_calc_range.argtypes = [C.c_void_p, _fp, _fp, C.c_int]
        lib.ref_numpy_calc_range_angles.argtypes = [C.c_void_p, _fp, _fp, _fp, C.c_int, C.c_int]
        lib.ref_eval_sensor_model.argtypes = [C.c_void_p, _fp, _fp, _dp, C.c_int, C.c_int]
        lib.ref_repeat_angles_eval_sensor_model.argtypes = [C.c_void_p, _fp, _fp, _fp, _dp, C.c_int, C.c_int]
        lib.ref_pf_step_mt.argtypes = [C.c_void_p, _fp, _fp, _fp, _fp, _dp, C.c_int, C.c_int, C.c_int]
        lib.ref_numpy_calc_range_mt.argtypes = [C.c_void_p, _fp, _fp, C.c_int, C.c_int]
        _REF = lib
    return _REF


def ref_load_png(path: str, threshold: float = 128.0) -> np.ndarray:
    """OMap(path, threshold) occupancy as an image ``img[y, x]``."""
    lib = ref_lib()
    w, h = C.c_int(), C.c_int()
    if lib.ref_load_png(path.encode(), threshold, C.byref(w), C.byref(h), None) != 0:
        raise IOError(path)
    occ = np.empty((w.value, h.value), dtype=np.uint8)
    lib.ref_load_png(path.encode(), threshold, C.byref(w), C.byref(h), _ptr(occ, _u8p))
    return np.ascontiguousarray(occ.T)


def ref_edt(img: np.ndarray) -> np.ndarray:
    occ = xmajor(img)
    w, h = occ.shape
    out = np.empty((w, h), dtype=np.float32)
    ref_lib().ref_edt(_ptr(occ, _u8p), w, h, _ptr(out, _fp))
    return out


class Reference:
    """The reference's RayMarching / BresenhamsLine on one map."""

    def __init__(self, img: np.ndarray, max_range: float, kind: str = "rm", world=None, theta_disc: int = 0):
        self.lib = ref_lib()
        occ = xmajor(img)
        w, h = occ.shape
        self.shape = (w, h, theta_disc)
        self._world = None if world is None else np.asarray(world, dtype=np.float32)
        wp = None if self._world is None else _ptr(self._world, _fp)
        if kind == "glt":
            self.h = self.lib.ref_create_glt(_ptr(occ, _u8p), w, h, max_range, theta_disc, wp)
        elif kind == "cddt":
            self.h = self.lib.ref_create_cddt(_ptr(occ, _u8p), w, h, max_range, theta_disc, wp)
        else:
            self.h = self.lib.ref_create(1 if kind == "bl" else 0, _ptr(occ, _u8p), w, h, max_range, wp)

    def calc_range_many_radial_optimized(self, ins, num_rays, min_angle, max_angle, fill=-7.0):
        """outs has one extra particle's slots of headroom: the reference writes pair
        results past the last particle's segment for some angle ranges."""
        ins = _f32(ins)
        outs = np.full((ins.shape[0] + 1) * num_rays, fill, dtype=np.float32)
        self.lib.ref_calc_range_many_radial_optimized(self.h, _ptr(ins, _fp), _ptr(outs, _fp), ins.shape[0],
                                                      num_rays, min_angle, max_angle)
        return outs[: ins.shape[0] * num_rays]

    def cddt_json(self) -> str:
        n = self.lib.ref_cddt_serialize_json(self.h, None, 0)
        buf = C.create_string_buffer(n + 1)
        self.lib.ref_cddt_serialize_json(self.h, buf, n + 1)
        return buf.raw[:n].decode()

    def cddt_prune(self, max_range: float):
        self.lib.ref_cddt_prune(self.h, max_range)

    def cddt_tables(self) -> dict:
        nb, nv = C.c_longlong(), C.c_longlong()
        self.lib.ref_cddt_sizes(self.h, C.byref(nb), C.byref(nv))
        td = self.shape[2]
        widths = np.empty(td, np.int32)
        trans = np.empty(td, np.float32)
        start = np.empty(nb.value + 1, np.int64)
        vals = np.empty(max(nv.value, 1), np.float32)
        self.lib.ref_cddt_export(self.h, widths.ctypes.data_as(C.POINTER(C.c_int)), _ptr(trans, _fp),
                                 start.ctypes.data_as(_llp), _ptr(vals, _fp))
        return {"widths": widths, "trans": trans, "start": start, "vals": vals[: nv.value]}

    def glt_table(self):
        """The reference's GiantLUTCast table (W, H, td) uint16 and (max_div_limits, limits_div_max)."""
        out = np.empty(self.shape, dtype=np.uint16)
        consts = np.empty(2, dtype=np.float32)
        self.lib.ref_glt_lut(self.h, out.ctypes.data_as(C.POINTER(C.c_uint16)), _ptr(consts, _fp))
        return out, consts

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.ref_destroy(self.h)
            self.h = None

    def set_sensor_model(self, table):
        self.table = np.ascontiguousarray(table, dtype=np.float64)
        self.lib.ref_set_sensor_model(self.h, _ptr(self.table, _dp), self.table.shape[0])

    def calc_range(self, x, y, heading):
        return self.lib.ref_calc_range(self.h, x, y, heading)

    def calc_range_many(self, ins):
        ins = _f32(ins)
        outs = np.empty(ins.shape[0], dtype=np.float32)
        self.lib.ref_calc_range_many(self.h, _ptr(ins, _fp), _ptr(outs, _fp), ins.shape[0])
        return outs

    def numpy_calc_range(self, ins, threads: int = 1):
        ins = _f32(ins)
        outs = np.empty(ins.shape[0], dtype=np.float32)
        if threads > 1:
            self.lib.ref_numpy_calc_range_mt(self.h, _ptr(ins, _fp), _ptr(outs, _fp), ins.shape[0], threads)
        else:
            self.lib.ref_numpy_calc_range(self.h, _ptr(ins, _fp), _ptr(outs, _fp), ins.shape[0])
        return outs

    def numpy_calc_range_angles(self, ins, angles):
        ins, angles = _f32(ins), _f32(angles)
        n, m = ins.shape[0], angles.shape[0]
        outs = np.empty(n * m, dtype=np.float32)
        self.lib.ref_numpy_calc_range_angles(self.h, _ptr(ins, _fp), _ptr(angles, _fp), _ptr(outs, _fp), n, m)
        return outs

    def eval_sensor_model(self, obs, ranges, num_rays, num_particles):
        obs, ranges = _f32(obs), _f32(ranges)
        outs = np.empty(num_particles, dtype=np.float64)
        self.lib.ref_eval_sensor_model(self.h, _ptr(obs, _fp), _ptr(ranges, _fp), _ptr(outs, _dp),
                                       num_rays, num_particles)
        return outs

    def calc_range_repeat_angles_eval_sensor_model(self, ins, angles, obs):
        ins, angles, obs = _f32(ins), _f32(angles), _f32(obs)
        n, m = ins.shape[0], angles.shape[0]
        w = np.empty(n, dtype=np.float64)
        self.lib.ref_repeat_angles_eval_sensor_model(self.h, _ptr(ins, _fp), _ptr(angles, _fp), _ptr(obs, _fp),
                                                     _ptr(w, _dp), n, m)
        return w

    def pf_step(self, ins, angles, obs, threads: int):
        """numpy_calc_range_angles + eval_sensor_model split over threads."""
        ins, angles, obs = _f32(ins), _f32(angles), _f32(obs)
        n, m = ins.shape[0], angles.shape[0]
        outs = np.empty(n * m, dtype=np.float32)
        w = np.empty(n, dtype=np.float64)
        self.lib.ref_pf_step_mt(self.h, _ptr(ins, _fp), _ptr(angles, _fp), _ptr(obs, _fp), _ptr(outs, _fp),
                                _ptr(w, _dp), n, m, threads)
        return outs, w
