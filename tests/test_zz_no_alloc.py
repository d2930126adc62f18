"""The no-allocation fused entry point against the allocating one, through the
C ABI: two calls in a row (the second smaller, so stale staging contents would
show), every weight bit for bit; sizes above the staging limit are refused."""
import ctypes as C

import numpy as np
import pytest

import range_libc_amd as rl
from range_libc_amd import _lib as L
from range_libc_amd import workloads as W
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


RL_ERR_ARG = 1  # include/rangelib_hip.h


def _err():
    msg = L.lib().rl_last_error()
    return msg.decode() if msg else ""


def _call(name, dm, R, ins, ang, obs):
    n, m = ins.shape[0], ang.shape[0]
    w = np.full(n, -1.0, np.float64)
    rc = getattr(L.lib(), name)(dm.handle, L.RL_METHOD_RM, C.c_float(R), _ptr(ins), _ptr(ang), _ptr(obs),
                                _ptr(w), n, m)
    return rc, w


def test_noalloc_fused_matches_allocating_call():
    ctx = rl.Context(0)
    img = W.corridor_map(1024, seed=7)
    dm = rl.DeviceMap(ctx, O.xmajor(img))
    table = W.sensor_table(501)
    dm.set_sensor_table(table)
    R = 500.0
    orc = O.Oracle(img, R, "rm")
    orc.set_sensor_model(table)
    for n, m, seed in ((20000, 60, 1), (1500, 30, 2)):
        ins = W.world_poses(img, n, seed=seed)
        ang = W.beam_angles(m)
        obs = W.noisy_observation(orc.numpy_calc_range_angles(W.world_poses(img, 1, seed=seed + 10), ang),
                                  seed=seed, max_range=R)
        rc0, want = _call("rl_calc_range_repeat_angles_eval_sensor_model", dm, R, ins, ang, obs)
        rc1, got = _call("rl_calc_range_repeat_angles_eval_sensor_model_noalloc", dm, R, ins, ang, obs)
        assert rc0 == 0 and rc1 == 0, _err()
        assert np.array_equal(got.view(np.uint64), want.view(np.uint64)), (n, m)
        if n <= 1500:
            assert np.array_equal(got, orc.calc_range_repeat_angles_eval_sensor_model(ins, ang, obs))
    # above the staging limit: refused, the output untouched
    big = W.world_poses(img, 300000, seed=5)
    ang = W.beam_angles(60)
    rc, w = _call("rl_calc_range_repeat_angles_eval_sensor_model_noalloc", dm, R, big, ang,
                  np.zeros(60, np.float32))
    assert rc == RL_ERR_ARG and "no-allocation" in _err()
    assert np.all(w == -1.0)
    dm.close()
