"""GPU results against the compiled reference's golden vectors (tests/golden/)."""
import os

import numpy as np
import pytest

import range_libc_amd as rl
from range_libc_amd import workloads as W
from tests.conftest import GOLDEN, assert_bits_equal
from tests.test_oracle import PNG_KEYS, npz, unpack

pytestmark = pytest.mark.gpu

CLS = {"rm": rl.PyRayMarchingGPU, "bl": rl.PyBresenhamsLine}


def _grid(meth, q):
    out = np.empty(q.shape[0], np.float32)
    meth.calc_range_many_grid(np.ascontiguousarray(q), out)
    return out


def _world(meth, q):
    out = np.empty(q.shape[0], np.float32)
    meth.calc_range_many(np.ascontiguousarray(q), out)
    return out


def test_gpu_edt_golden():
    z = npz("png_maps.npz")
    ctx = rl.default_context(0)
    for key in PNG_KEYS:
        img = unpack(z, f"{key}_t128")
        got = rl.DeviceMap(ctx, np.ascontiguousarray(img.T)).edt()
        if f"{key}_edt" in z:
            assert_bits_equal(got, z[f"{key}_edt"], key)
        else:
            assert_bits_equal(got.reshape(-1)[z[f"{key}_edt_idx"]], z[f"{key}_edt_val"], key)
    zd = npz("degenerate.npz")
    for name in ("empty", "full", "one_free", "checker", "tall"):
        got = rl.DeviceMap(ctx, np.ascontiguousarray(zd[f"{name}_img"].T)).edt()
        assert_bits_equal(got, zd[f"{name}_edt"], name)


def test_gpu_ranges_golden_png_maps():
    z = npz("png_maps.npz")
    for key in PNG_KEYS:
        omap = rl.PyOMap(unpack(z, f"{key}_t128"))
        for kind in ("rm", "bl"):
            m = CLS[kind](omap, 500.0)
            eq = z[f"{key}_{kind}_edge_ins"]
            assert_bits_equal(_grid(m, eq), z[f"{key}_{kind}_edge_grid"], f"{key} {kind} edge grid")
            assert_bits_equal(_world(m, eq), z[f"{key}_{kind}_edge_world"], f"{key} {kind} edge world")
            if f"{key}_{kind}_q" in z:
                assert_bits_equal(_world(m, z[f"{key}_{kind}_q"]), z[f"{key}_{kind}_world"], key)


def test_gpu_ranges_golden_degenerate():
    z = npz("degenerate.npz")
    for name in ("empty", "full", "one_free", "checker", "tall"):
        omap = rl.PyOMap(z[f"{name}_img"])
        eq = z[f"{name}_ins"]
        for kind in ("rm", "bl"):
            for mr in (0, 1, 50):
                m = CLS[kind](omap, float(mr))
                assert_bits_equal(_grid(m, eq), z[f"{name}_{kind}_{mr}_grid"], f"{name} {kind} {mr}")
                assert_bits_equal(_world(m, eq), z[f"{name}_{kind}_{mr}_world"], f"{name} {kind} {mr}")


@pytest.mark.parametrize("kind", ["rm", "bl"])
def test_gpu_golden_rect_disk_512(kind):
    z = npz("rect_disk_512.npz")
    img = W.rect_disk_map(512, seed=0)
    m = CLS[kind](rl.PyOMap(img), 500.0)
    m.set_sensor_model(W.sensor_table(501))
    assert_bits_equal(_world(m, z[f"{kind}_qw"]), z[f"{kind}_world"], "world")
    assert_bits_equal(_grid(m, z[f"{kind}_qg"]), z[f"{kind}_grid"], "grid")
    ang = W.beam_angles(60)
    parts = z[f"{kind}_parts"]
    rr = np.empty(50 * 60, np.float32)
    m.calc_range_repeat_angles(parts, ang, rr)
    assert_bits_equal(rr, z[f"{kind}_angles"], "angles")
    w = np.empty(50)
    m.eval_sensor_model(z[f"{kind}_obs"], rr, w, 60, 50)
    assert np.array_equal(w, z[f"{kind}_sensor"])
    m.calc_range_repeat_angles_eval_sensor_model(parts, ang, z[f"{kind}_obs"], w)
    assert np.array_equal(w, z[f"{kind}_fused"])
    omap = rl.PyOMap(img)
    omap.set_world(*[float(v) for v in z["world"]])
    mw = CLS[kind](omap, 500.0)
    assert_bits_equal(_world(mw, z[f"{kind}_wq"]), z[f"{kind}_wworld"], "world params")


@pytest.mark.parametrize("kind", ["rm", "bl"])
def test_gpu_golden_corridor_2048(kind):
    z = npz("corridor_2048.npz")
    img = W.corridor_map(2048, seed=1)
    omap = rl.PyOMap(img)
    rr = np.empty(200 * 60, np.float32)
    CLS[kind](omap, 500.0).calc_range_repeat_angles(z["parts"], W.beam_angles(60), rr)
    assert_bits_equal(rr, z[f"{kind}_angles"], "angles")
    assert_bits_equal(_world(CLS[kind](omap, 1000.0), z[f"{kind}_qi"]), z[f"{kind}_r1000"], "r1000")
