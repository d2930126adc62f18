"""ctypes binding of librangelib_hip.so (the C ABI in include/rangelib_hip.h).

Loading the shared library needs no GPU; creating a context does.  There is
no CPU fallback: if the library is missing, every entry point raises.
"""
from __future__ import annotations

import ctypes as C
import os
import threading

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("RL_LIBRARY", os.path.join(PKG, "librangelib_hip.so"))

RL_METHOD_RM = 0
RL_METHOD_BL = 1
RL_METHOD_GLT = 2
RL_METHOD_CDDT = 3
RL_FRAME_GRID = 0
RL_FRAME_WORLD = 1
RL_FRAME_ANGLES = 2
RL_MAP_EDT_ENVELOPE_ONLY = 1
RL_MAP_NO_COMPACT = 2
RL_OPT_XCD_SCHED = 1
RL_OPT_LDS_BOX = 2

_vp = C.c_void_p
_fp = C.POINTER(C.c_float)
_dp = C.POINTER(C.c_double)
_u8p = C.POINTER(C.c_uint8)

# name -> argument types; every function returns int status
SIGNATURES = {
    "rl_abi_version": [],
    "rl_trig_table_info": [C.c_char_p, C.c_char_p, C.c_int],
    "rl_png_occupancy": [C.c_char_p, C.c_float, _vp, C.POINTER(C.c_int), C.POINTER(C.c_int)],
    "rl_png_save_occupancy": [C.c_char_p, _vp, C.c_int, C.c_int],
    "rl_create": [C.c_int, C.POINTER(_vp)],
    "rl_destroy": [_vp],
    "rl_set_stream": [_vp, _vp],
    "rl_sync": [_vp],
    "rl_set_option": [_vp, C.c_int, C.c_longlong],
    "rl_map_create": [_vp, _u8p, C.c_int, C.c_int, C.POINTER(_vp)],
    "rl_map_create_ex": [_vp, _u8p, C.c_int, C.c_int, C.c_uint, C.POINTER(_vp)],
    "rl_map_create_from_edt": [_vp, _u8p, _fp, C.c_int, C.c_int, C.c_uint, C.POINTER(_vp)],
    "rl_map_destroy": [_vp],
    "rl_map_build_giant_lut": [_vp, C.c_float, C.c_int],
    "rl_map_build_cddt": [_vp, C.c_int],
    "rl_calc_range_many_radial_optimized": [_vp, C.c_int, C.c_float, _vp, _vp, C.c_int, C.c_int, C.c_float, C.c_float],
    "rl_calc_range_many_radial_optimized_dev": [_vp, C.c_int, C.c_float, _vp, _vp, C.c_int, C.c_int, C.c_float,
                                                C.c_float],
    "rl_map_prune_cddt": [_vp, C.c_float],
    "rl_map_get_cddt": [_vp, C.POINTER(C.c_int), C.POINTER(C.c_longlong), C.POINTER(C.c_longlong), _vp, _vp, _vp,
                        _vp, _fp],
    "rl_map_get_giant_lut": [_vp, _vp, C.POINTER(C.c_int), _fp, _fp],
    "rl_map_set_world": [_vp] + [C.c_float] * 6,
    "rl_map_set_sensor_table": [_vp, _dp, C.c_int],
    "rl_map_get_edt": [_vp, _fp],
    "rl_map_device_buffers": [_vp, C.POINTER(_vp), C.POINTER(_vp)],
    "rl_map_edt_build_ms": [_vp, _fp],
    "rl_calc_range": [_vp, C.c_int, C.c_float, C.c_float, C.c_float, C.c_float, _fp],
    "rl_calc_range_many": [_vp, C.c_int, C.c_float, _vp, _vp, C.c_int],
    "rl_numpy_calc_range": [_vp, C.c_int, C.c_float, _vp, _vp, C.c_int],
    "rl_numpy_calc_range_angles": [_vp, C.c_int, C.c_float, _vp, _vp, _vp, C.c_int, C.c_int],
    "rl_eval_sensor_model": [_vp, _vp, _vp, _vp, C.c_int, C.c_int],
    "rl_calc_range_repeat_angles_eval_sensor_model": [_vp, C.c_int, C.c_float, _vp, _vp, _vp, _vp,
                                                      C.c_int, C.c_int],
    "rl_calc_range_repeat_angles_eval_sensor_model_noalloc": [_vp, C.c_int, C.c_float, _vp, _vp, _vp, _vp,
                                                              C.c_int, C.c_int],
    "rl_calc_range_many_dev": [_vp, C.c_int, C.c_float, _vp, _vp, C.c_int],
    "rl_numpy_calc_range_dev": [_vp, C.c_int, C.c_float, _vp, _vp, C.c_int],
    "rl_numpy_calc_range_angles_dev": [_vp, C.c_int, C.c_float, _vp, _vp, _vp, C.c_int, C.c_int],
    "rl_eval_sensor_model_dev": [_vp, _vp, _vp, _vp, C.c_int, C.c_int],
    "rl_calc_range_repeat_angles_eval_sensor_model_dev": [_vp, C.c_int, C.c_float, _vp, _vp, _vp,
                                                          _vp, C.c_int, C.c_int],
    "rl_pf_step_dev": [_vp, C.c_int, C.c_float, _vp, _vp, _vp, _vp, _vp, C.c_int, C.c_int],
    "rl_count_steps_dev": [_vp, C.c_int, C.c_int, C.c_float, _vp, _vp, _vp, C.c_int, C.c_int],
    "rl_batch_create": [_vp, C.c_int, C.POINTER(_vp)],
    "rl_batch_destroy": [_vp],
    "rl_batch_set_jobs": [_vp, _vp, C.c_int],
    "rl_batch_autotune": [_vp, C.c_int, _fp],
    "rl_batch_set_config": [_vp, C.c_int, C.c_int, C.c_int, C.c_int],
    "rl_batch_get_config": [_vp, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)],
    "rl_batch_launch": [_vp],
    "rl_debug_sincosf_dev": [_vp, _vp, _vp, _vp, C.c_int],
    "rl_debug_sqrtf_dev": [_vp, _vp, _vp, C.c_int],
    "rl_debug_chase_dev": [_vp, _vp, C.c_int, _vp, C.POINTER(C.c_double)],
    "rl_debug_exclusive_sum_dev": [_vp, _vp, _vp, C.c_longlong],
    "rl_debug_segmented_sort_dev": [_vp, _vp, _vp, _vp, C.c_longlong],
    "rl_debug_gemm_bf16": [_vp, _vp, _vp, _vp, C.c_int, C.c_int, C.c_int],
    "rl_debug_gemm_bf16_nt": [_vp, _vp, _vp, _vp, C.c_int, C.c_int, C.c_int],
    "rl_surrogate_create": [_vp, C.c_uint, C.POINTER(_vp)],
    "rl_surrogate_destroy": [_vp],
    "rl_surrogate_num_params": [],
    "rl_surrogate_get_params": [_vp, _vp, _vp],
    "rl_surrogate_set_params": [_vp, _vp, _vp],
    "rl_surrogate_forward": [_vp, _vp, C.c_int, _vp],
    "rl_debug_surrogate_forward_layers": [_vp, _vp, C.c_int, _vp],
    "rl_debug_bf16_check": [_vp, _vp],
    "rl_surrogate_gradient": [_vp, C.c_float, _vp, _vp, C.c_int, _fp, _vp, _vp],
    "rl_surrogate_step": [_vp, C.c_float],
    "rl_surrogate_train": [_vp, C.c_float, C.c_longlong, C.c_int, C.c_float, C.c_uint, _vp],
    "rl_surrogate_calc_range_angles_dev": [_vp, C.c_float, _vp, _vp, _vp, C.c_int, C.c_int, C.c_float,
                                           C.POINTER(C.c_uint)],
}


class RangeLibError(RuntimeError):
    """A call into librangelib_hip failed; the message is rl_last_error()."""

    def __init__(self, code: int, message: str):
        super().__init__(f"[rl error {code}] {message}")
        self.code = code


_LIB = None
_LOCK = threading.Lock()


def lib():
    """The loaded library; raises if it has not been built."""
    global _LIB
    with _LOCK:
        if _LIB is None:
            if not os.path.exists(LIB_PATH):
                raise RangeLibError(-1, f"HIP library not built: {LIB_PATH} "
                                        "(run `python -m range_libc_amd.build`)")
            # One HIP runtime per process: PyTorch ships its own libamdhip64 /
            # libhsa-runtime64 and links them by unversioned name, so if this
            # library loaded the system copies first, a later `import torch`
            # would bring in a second runtime.  Loading torch first makes this
            # library bind to torch's runtime (same sonames).
            try:
                import torch  # noqa: F401
            except ImportError:
                pass
            h = C.CDLL(LIB_PATH)
            for name, args in SIGNATURES.items():
                fn = getattr(h, name)
                fn.argtypes = args
                fn.restype = C.c_int
            h.rl_last_error.argtypes = []
            h.rl_last_error.restype = C.c_char_p
            _LIB = h
    return _LIB


def check(rc: int):
    if rc != 0:
        msg = lib().rl_last_error()
        raise RangeLibError(rc, msg.decode() if msg else "unknown error")


def ptr(a: np.ndarray):
    return C.c_void_p(a.ctypes.data)


class Context:
    """One HIP device and stream (rl_ctx)."""

    def __init__(self, device: int = 0):
        h = C.c_void_p()
        check(lib().rl_create(device, C.byref(h)))
        self.handle = h
        self.device = device

    def close(self):
        if getattr(self, "handle", None):
            lib().rl_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_stream(self, stream_handle: int | None):
        """Order work on a raw hipStream_t (e.g. torch.cuda.current_stream().cuda_stream)."""
        check(lib().rl_set_stream(self.handle, C.c_void_p(stream_handle or 0)))

    def sync(self):
        check(lib().rl_sync(self.handle))

    def set_option(self, option: int, value: int):
        check(lib().rl_set_option(self.handle, option, int(value)))

    def sincosf_dev(self, x_ptr: int, s_ptr: int, c_ptr: int, n: int):
        check(lib().rl_debug_sincosf_dev(self.handle, _vp(x_ptr), _vp(s_ptr), _vp(c_ptr), n))

    def sqrtf_dev(self, x_ptr: int, out_ptr: int, n: int):
        check(lib().rl_debug_sqrtf_dev(self.handle, _vp(x_ptr), _vp(out_ptr), n))

    def debug_chase(self, next_ptr: int, steps: int, out_ptr: int) -> float:
        """Dependent-load latency probe; returns ns per wall-clock tick."""
        tick = C.c_double()
        check(lib().rl_debug_chase_dev(self.handle, _vp(next_ptr), steps, _vp(out_ptr), C.byref(tick)))
        return tick.value

    def exclusive_sum_dev(self, in_ptr: int, out_ptr: int, n: int):
        """The CDDT build's device-wide exclusive sum of uint64 counts."""
        check(lib().rl_debug_exclusive_sum_dev(self.handle, _vp(in_ptr), _vp(out_ptr), n))

    def segmented_sort_dev(self, in_ptr: int, out_ptr: int, start_ptr: int, nseg: int):
        """The CDDT build's per-segment ascending float sort (int64 offsets)."""
        check(lib().rl_debug_segmented_sort_dev(self.handle, _vp(in_ptr), _vp(out_ptr), _vp(start_ptr), nseg))


_DEFAULT: dict[int, Context] = {}


def default_context(device: int = 0) -> Context:
    with _LOCK:
        ctx = _DEFAULT.get(device)
    if ctx is None:
        ctx = Context(device)
        with _LOCK:
            _DEFAULT.setdefault(device, ctx)
            ctx = _DEFAULT[device]
    return ctx


class DeviceMap:
    """An occupancy grid resident on the GPU with its distance transform (rl_map).

    ``occ_xmajor`` is uint8 (width, height): ``occ[x, y]`` = grid[x][y].
    """

    def __init__(self, ctx: Context, occ_xmajor: np.ndarray, flags: int = 0, edt_xmajor: np.ndarray | None = None):
        occ = np.ascontiguousarray(occ_xmajor, dtype=np.uint8)
        if occ.ndim != 2:
            raise ValueError("occupancy must be 2-D (width, height)")
        self.ctx = ctx
        self.width, self.height = int(occ.shape[0]), int(occ.shape[1])
        h = C.c_void_p()
        if edt_xmajor is None:
            check(lib().rl_map_create_ex(ctx.handle, occ.ctypes.data_as(_u8p), self.width, self.height,
                                         flags, C.byref(h)))
        else:
            edt = np.ascontiguousarray(edt_xmajor, dtype=np.float32)
            if edt.shape != occ.shape:
                raise ValueError("distance transform and occupancy shapes differ")
            check(lib().rl_map_create_from_edt(ctx.handle, occ.ctypes.data_as(_u8p), edt.ctypes.data_as(_fp),
                                               self.width, self.height, flags, C.byref(h)))
        self.handle = h
        self.has_table = False

    def close(self):
        if getattr(self, "handle", None):
            lib().rl_map_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- state ---------------------------------------------------------------
    def set_world(self, scale, angle, origin_x, origin_y, sin_angle, cos_angle):
        check(lib().rl_map_set_world(self.handle, scale, angle, origin_x, origin_y, sin_angle, cos_angle))

    def set_sensor_table(self, table: np.ndarray):
        t = np.ascontiguousarray(table, dtype=np.float64)
        check(lib().rl_map_set_sensor_table(self.handle, t.ctypes.data_as(_dp), t.shape[0]))
        self.has_table = True

    def build_giant_lut(self, max_range: float, theta_disc: int):
        """GiantLUTCast table of this map (RL_METHOD_GLT then works at this max_range)."""
        check(lib().rl_map_build_giant_lut(self.handle, max_range, int(theta_disc)))

    def giant_lut(self, copy: bool = True):
        """(table (W, H, td) uint16 or None, td, max_range, build_ms)."""
        td, r, ms = C.c_int(), C.c_float(), C.c_float()
        check(lib().rl_map_get_giant_lut(self.handle, None, C.byref(td), C.byref(r), C.byref(ms)))
        out = None
        if copy:
            out = np.empty((self.width, self.height, td.value), dtype=np.uint16)
            check(lib().rl_map_get_giant_lut(self.handle, C.c_void_p(out.ctypes.data), None, None, None))
        return out, td.value, r.value, ms.value

    def calc_range_many_radial_optimized(self, method, max_range, ins, outs, num_particles, num_rays, min_angle,
                                         max_angle):
        check(lib().rl_calc_range_many_radial_optimized(self.handle, method, max_range, ptr(ins), ptr(outs),
                                                        num_particles, num_rays, min_angle, max_angle))

    def build_cddt(self, theta_disc: int):
        """CDDTCast tables of this map (RL_METHOD_CDDT then works)."""
        check(lib().rl_map_build_cddt(self.handle, int(theta_disc)))

    def prune_cddt(self, max_range: float):
        """CDDTCast::prune: keep only the entries a query can return (PCDDT)."""
        check(lib().rl_map_prune_cddt(self.handle, max_range))

    def cddt_tables(self) -> dict:
        """CSR copy of the tables: widths, trans, start, vals, build_ms."""
        td, nb, nv, ms = C.c_int(), C.c_longlong(), C.c_longlong(), C.c_float()
        check(lib().rl_map_get_cddt(self.handle, C.byref(td), C.byref(nb), C.byref(nv), None, None, None, None,
                                    C.byref(ms)))
        widths = np.empty(td.value, np.int32)
        trans = np.empty(td.value, np.float32)
        start = np.empty(nb.value + 1, np.int64)
        vals = np.empty(max(nv.value, 1), np.float32)
        check(lib().rl_map_get_cddt(self.handle, None, None, None, C.c_void_p(widths.ctypes.data),
                                    C.c_void_p(trans.ctypes.data), C.c_void_p(start.ctypes.data),
                                    C.c_void_p(vals.ctypes.data), None))
        return {"widths": widths, "trans": trans, "start": start, "vals": vals[: nv.value], "build_ms": ms.value}

    def edt(self) -> np.ndarray:
        out = np.empty((self.width, self.height), dtype=np.float32)
        check(lib().rl_map_get_edt(self.handle, out.ctypes.data_as(_fp)))
        return out

    def edt_build_ms(self) -> float:
        v = C.c_float()
        check(lib().rl_map_edt_build_ms(self.handle, C.byref(v)))
        return v.value

    def device_buffers(self):
        e, o = C.c_void_p(), C.c_void_p()
        check(lib().rl_map_device_buffers(self.handle, C.byref(e), C.byref(o)))
        return e.value, o.value

    # -- host-pointer batches (synchronous) ------------------------------------
    def calc_range(self, method: int, max_range: float, x: float, y: float, heading: float) -> float:
        out = C.c_float()
        check(lib().rl_calc_range(self.handle, method, max_range, x, y, heading, C.byref(out)))
        return out.value

    def calc_range_many(self, method, max_range, ins, outs, n):
        check(lib().rl_calc_range_many(self.handle, method, max_range, ptr(ins), ptr(outs), n))

    def numpy_calc_range(self, method, max_range, ins, outs, n):
        check(lib().rl_numpy_calc_range(self.handle, method, max_range, ptr(ins), ptr(outs), n))

    def numpy_calc_range_angles(self, method, max_range, ins, angles, outs, n, m):
        check(lib().rl_numpy_calc_range_angles(self.handle, method, max_range, ptr(ins), ptr(angles),
                                               ptr(outs), n, m))

    def eval_sensor_model(self, obs, ranges, outs, num_rays, num_particles):
        check(lib().rl_eval_sensor_model(self.handle, ptr(obs), ptr(ranges), ptr(outs), num_rays,
                                         num_particles))

    def repeat_angles_eval_sensor_model(self, method, max_range, ins, angles, obs, weights, n, m):
        check(lib().rl_calc_range_repeat_angles_eval_sensor_model(
            self.handle, method, max_range, ptr(ins), ptr(angles), ptr(obs), ptr(weights), n, m))

    # -- device-pointer batches (stream-ordered, integer addresses) -------------
    def calc_range_many_dev(self, method, max_range, ins, outs, n):
        check(lib().rl_calc_range_many_dev(self.handle, method, max_range, _vp(ins), _vp(outs), n))

    def numpy_calc_range_dev(self, method, max_range, ins, outs, n):
        check(lib().rl_numpy_calc_range_dev(self.handle, method, max_range, _vp(ins), _vp(outs), n))

    def numpy_calc_range_angles_dev(self, method, max_range, ins, angles, outs, n, m):
        check(lib().rl_numpy_calc_range_angles_dev(self.handle, method, max_range, _vp(ins), _vp(angles),
                                                   _vp(outs), n, m))

    def eval_sensor_model_dev(self, obs, ranges, outs, num_rays, num_particles):
        check(lib().rl_eval_sensor_model_dev(self.handle, _vp(obs), _vp(ranges), _vp(outs), num_rays,
                                             num_particles))

    def repeat_angles_eval_sensor_model_dev(self, method, max_range, ins, angles, obs, weights, n, m):
        check(lib().rl_calc_range_repeat_angles_eval_sensor_model_dev(
            self.handle, method, max_range, _vp(ins), _vp(angles), _vp(obs), _vp(weights), n, m))

    def pf_step_dev(self, method, max_range, ins, angles, obs, ranges, weights, n, m):
        check(lib().rl_pf_step_dev(self.handle, method, max_range, _vp(ins), _vp(angles), _vp(obs), _vp(ranges),
                                   _vp(weights), n, m))

    def count_steps_dev(self, method, frame, max_range, ins, angles, counts, n, m=0):
        check(lib().rl_count_steps_dev(self.handle, method, frame, max_range, _vp(ins), _vp(angles or 0),
                                       _vp(counts), n, m))


class BatchJob(C.Structure):
    """rl_batch_job: one map's particle-filter work inside a multi-map batch."""
    _fields_ = [("map", _vp), ("max_range", C.c_float), ("d_ins", _vp), ("d_angles", _vp), ("d_obs", _vp),
                ("d_ranges", _vp), ("d_weights", _vp), ("num_particles", C.c_int), ("num_angles", C.c_int)]


class Batch:
    """Many maps cast in one launch (rl_batch).  Jobs hold device addresses."""

    def __init__(self, ctx: Context, method: int = RL_METHOD_RM):
        h = C.c_void_p()
        check(lib().rl_batch_create(ctx.handle, method, C.byref(h)))
        self.handle = h
        self.ctx = ctx
        self._keep = []

    def close(self):
        if getattr(self, "handle", None):
            lib().rl_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_jobs(self, jobs):
        """jobs: list of dicts with keys map (DeviceMap), max_range, ins, angles, obs (or 0),
        ranges, weights (or 0), num_particles, num_angles (device addresses as ints)."""
        arr = (BatchJob * len(jobs))()
        for i, j in enumerate(jobs):
            arr[i] = BatchJob(j["map"].handle, j["max_range"], j["ins"], j["angles"], j.get("obs") or None,
                              j["ranges"], j.get("weights") or None, j["num_particles"], j["num_angles"])
        self._keep = [j["map"] for j in jobs]
        check(lib().rl_batch_set_jobs(self.handle, C.cast(arr, _vp), len(jobs)))

    def autotune(self, reps: int = 3) -> float:
        ms = C.c_float()
        check(lib().rl_batch_autotune(self.handle, reps, C.byref(ms)))
        return ms.value

    def set_config(self, rays_per_wave=64, lds_box=0, order=1, grid=0):
        check(lib().rl_batch_set_config(self.handle, rays_per_wave, lds_box, order, grid))

    def config(self):
        v = [C.c_int() for _ in range(4)]
        check(lib().rl_batch_get_config(self.handle, *[C.byref(x) for x in v]))
        return dict(zip(("rays_per_wave", "lds_box", "order", "grid"), (x.value for x in v)))

    def launch(self):
        check(lib().rl_batch_launch(self.handle))


def debug_gemm_bf16(ctx: Context, a_bits: np.ndarray, b_bits: np.ndarray) -> np.ndarray:
    """A (M x K) @ B (K x N) of bf16 bit patterns (uint16) on the matrix cores; fp32 result."""
    a = np.ascontiguousarray(a_bits, dtype=np.uint16)
    b = np.ascontiguousarray(b_bits, dtype=np.uint16)
    M, K = a.shape
    K2, N = b.shape
    if K != K2:
        raise ValueError("inner dimensions differ")
    c = np.empty((M, N), dtype=np.float32)
    check(lib().rl_debug_gemm_bf16(ctx.handle, ptr(a), ptr(b), ptr(c), M, N, K))
    return c


def debug_bf16_check(ctx: Context) -> int:
    """Float bit patterns (of all 2^32) where the fused surrogate kernel's bf16
    conversion differs from the per-layer kernels'."""
    n = C.c_ulonglong()
    check(lib().rl_debug_bf16_check(ctx.handle, C.byref(n)))
    return int(n.value)


def debug_gemm_bf16_nt(ctx: Context, a_bits: np.ndarray, bt_bits: np.ndarray) -> np.ndarray:
    """A (M x K) @ Bt (N x K)^T of bf16 bit patterns; fp32 result (the forward-layer layout)."""
    a = np.ascontiguousarray(a_bits, dtype=np.uint16)
    bt = np.ascontiguousarray(bt_bits, dtype=np.uint16)
    M, K = a.shape
    N, K2 = bt.shape
    if K != K2:
        raise ValueError("inner dimensions differ")
    c = np.empty((M, N), dtype=np.float32)
    check(lib().rl_debug_gemm_bf16_nt(ctx.handle, ptr(a), ptr(bt), ptr(c), M, N, K))
    return c


class Surrogate:
    """Learned per-map range surrogate (rl_surrogate_*): train on the GPU, then
    cast particle batches through the checked start point."""

    NUM_FREQ = 32

    def __init__(self, dmap: "DeviceMap", seed: int = 0):
        h = C.c_void_p()
        check(lib().rl_surrogate_create(dmap.handle, seed, C.byref(h)))
        self.handle = h
        self.map = dmap
        self.num_params = int(lib().rl_surrogate_num_params())

    def close(self):
        if getattr(self, "handle", None):
            lib().rl_surrogate_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def params(self):
        p = np.empty(self.num_params, np.float32)
        f = np.empty((self.NUM_FREQ, 4), np.float32)
        check(lib().rl_surrogate_get_params(self.handle, ptr(p), ptr(f)))
        return p, f

    def set_params(self, params=None, freq=None):
        pp = None if params is None else np.ascontiguousarray(params, np.float32)
        ff = None if freq is None else np.ascontiguousarray(freq, np.float32)
        if pp is not None and pp.size != self.num_params:
            raise ValueError("wrong parameter count")
        if ff is not None and ff.size != self.NUM_FREQ * 4:
            raise ValueError("wrong frequency count")
        check(lib().rl_surrogate_set_params(self.handle, None if pp is None else ptr(pp),
                                            None if ff is None else ptr(ff)))

    def forward(self, rays4: np.ndarray) -> np.ndarray:
        r = np.ascontiguousarray(rays4, np.float32).reshape(-1, 4)
        out = np.empty((r.shape[0], 2), np.float32)
        check(lib().rl_surrogate_forward(self.handle, ptr(r), r.shape[0], ptr(out)))
        return out

    def forward_layers(self, rays4: np.ndarray) -> np.ndarray:
        """forward() through the per-layer kernels that training uses (test hook)."""
        r = np.ascontiguousarray(rays4, np.float32).reshape(-1, 4)
        out = np.empty((r.shape[0], 2), np.float32)
        check(lib().rl_debug_surrogate_forward_layers(self.handle, ptr(r), r.shape[0], ptr(out)))
        return out

    def gradient(self, max_range: float, rays4: np.ndarray, targets: np.ndarray):
        """(loss, grad, features): features are the bf16 bit patterns the network saw."""
        r = np.ascontiguousarray(rays4, np.float32).reshape(-1, 4)
        t = np.ascontiguousarray(targets, np.float32).reshape(-1)
        n = r.shape[0]
        if t.size != n:
            raise ValueError("one target per ray")
        g = np.empty(self.num_params, np.float32)
        feat = np.empty((n, 64), np.uint16)
        loss = C.c_float()
        check(lib().rl_surrogate_gradient(self.handle, max_range, ptr(r), ptr(t), n, C.byref(loss), ptr(g),
                                          ptr(feat)))
        return loss.value, g, feat

    def step(self, lr: float = 1e-3):
        """One Adam step with the gradient of the last ``gradient`` call."""
        check(lib().rl_surrogate_step(self.handle, lr))

    def train(self, max_range: float, n_pairs: int, batch: int = 65536, lr: float = 1e-3, seed: int = 1,
              curve: bool = True):
        steps = n_pairs // batch
        c = np.zeros(max(steps, 1), np.float32) if curve else None
        check(lib().rl_surrogate_train(self.handle, max_range, n_pairs, batch, lr, seed,
                                       None if c is None else ptr(c)))
        return c[:steps] if c is not None else None

    def calc_range_angles_dev(self, max_range, d_ins, d_angles, d_ranges, num_particles, num_angles,
                              bound_max=0.5, stats=True):
        st = (C.c_uint * 5)() if stats else None
        check(lib().rl_surrogate_calc_range_angles_dev(self.handle, max_range, _vp(d_ins), _vp(d_angles),
                                                       _vp(d_ranges), num_particles, num_angles, bound_max,
                                                       None if st is None else C.cast(st, C.POINTER(C.c_uint))))
        if st is None:
            return None
        return dict(zip(("fallback_bound", "fallback_check", "from_start", "reads", "rounds"), (int(v) for v in st)))

