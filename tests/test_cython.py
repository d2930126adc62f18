"""The Cython module `range_libc` (range_libc_amd/pywrapper/RangeLibc.pyx),
built over the C ABI: the reference's module surface (RangeLibc.pyx:130-351),
compiled and loaded without a GPU, and every casting method bit for bit
against the oracle on the GPU."""
import os
import subprocess
import sys
from types import SimpleNamespace as NS

import numpy as np
import pytest

from oracle import oracle as O
from range_libc_amd import workloads as W
from tests.conftest import GOLDEN, ROOT, assert_bits_equal

PYWRAPPER = os.path.join(ROOT, "range_libc_amd", "pywrapper")

# cpdef methods of each class in the reference module (RangeLibc.pyx:185-351)
REFERENCE_METHODS = {
    "PyOMap": {"save", "isOccupied", "error", "width", "height"},
    "PyBresenhamsLine": {"calc_range", "calc_range_many", "calc_range_repeat_angles",
                         "calc_range_repeat_angles_eval_sensor_model", "saveTrace", "eval_sensor_model",
                         "set_sensor_model"},
    "PyRayMarching": {"calc_range", "calc_range_many", "saveTrace", "calc_range_repeat_angles",
                      "calc_range_repeat_angles_eval_sensor_model", "eval_sensor_model", "set_sensor_model"},
    "PyCDDTCast": {"prune", "calc_range", "calc_range_many", "calc_range_repeat_angles_eval_sensor_model",
                   "calc_range_many_radial_optimized", "calc_range_repeat_angles", "eval_sensor_model",
                   "set_sensor_model"},
    "PyGiantLUTCast": {"calc_range", "calc_range_many", "calc_range_repeat_angles",
                       "calc_range_repeat_angles_eval_sensor_model", "eval_sensor_model", "set_sensor_model"},
    "PyRayMarchingGPU": {"calc_range_many", "calc_range_repeat_angles", "calc_range_repeat_angles_eval_sensor_model",
                         "eval_sensor_model", "set_sensor_model"},
    "PyNull": {"calc_range", "calc_range_many"},
}


@pytest.fixture(scope="module")
def range_libc():
    if PYWRAPPER not in sys.path:
        sys.path.insert(0, PYWRAPPER)
    import range_libc as mod

    return mod


def test_module_surface_matches_reference(range_libc):
    for cls, methods in REFERENCE_METHODS.items():
        c = getattr(range_libc, cls)
        have = {k for k in dir(c) if not k.startswith("_")}
        assert methods <= have, (cls, methods - have)
    assert range_libc.USE_ALTERNATE_MOD is True and range_libc.SHOULD_USE_CUDA is True


def test_omap_constructors_without_gpu(range_libc):
    z = np.load(os.path.join(GOLDEN, "png_maps.npz"))
    for key, name in (("synthetic_map", "synthetic.map.png"), ("basement_hallways_5cm", "basement_hallways_5cm.png")):
        for thr in (128, 1):
            shape = tuple(int(v) for v in z[f"{key}_t{thr}_shape"])
            want = np.unpackbits(z[f"{key}_t{thr}_occ"])[: shape[0] * shape[1]].reshape(shape)
            m = range_libc.PyOMap(os.path.join(GOLDEN, "maps", name), float(thr))
            assert not m.error() and (m.width(), m.height()) == (shape[1], shape[0])
            assert np.array_equal(m.occ.T, want)
    img = W.rect_disk_map(40, seed=1)[:, :30]
    m = range_libc.PyOMap(img)
    assert (m.width(), m.height()) == (30, 40) and m.isOccupied(0, 0) == bool(img[0, 0])
    e = range_libc.PyOMap(8, 5)
    assert (e.width(), e.height()) == (8, 5) and not e.isOccupied(3, 3)
    bad = range_libc.PyOMap(os.path.join(ROOT, "missing.png"))
    assert bad.error()
    h, w = 4, 6
    data = np.arange(h * w) % 3 * 50
    q = NS(x=0.0, y=0.0, z=np.sin(0.2), w=np.cos(0.2))
    msg = NS(info=NS(width=w, height=h, resolution=0.1, origin=NS(position=NS(x=1.0, y=2.0, z=0.0), orientation=q)),
             data=list(data))
    r = range_libc.PyOMap(msg)
    assert (r.width(), r.height()) == (h, w)
    assert np.array_equal(r.occ, (data.reshape(h, w) > 10).astype(np.uint8))
    assert abs(r.world_angle + 0.4) < 1e-6 and abs(r.world_scale - 0.1) < 1e-7


def test_save_round_trip(range_libc, tmp_path):
    m = range_libc.PyOMap(W.rect_disk_map(32, seed=2))
    p = str(tmp_path / "m.png")
    assert m.save(p) is False
    assert np.array_equal(range_libc.PyOMap(p).occ, m.occ)


def test_methods_fail_loudly_without_device():
    code = ("import sys; sys.path.insert(0, %r); import range_libc, numpy as np\n"
            "m = range_libc.PyOMap(np.zeros((8, 8), np.uint8))\n"
            "try:\n    range_libc.PyRayMarchingGPU(m, 10.0)\nexcept RuntimeError as e:\n    print('raised', e)\n"
            % PYWRAPPER)
    env = dict(os.environ, HIP_VISIBLE_DEVICES="-1", ROCR_VISIBLE_DEVICES="-1")
    out = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, env=env, timeout=120).stdout
    assert out.startswith("raised rangelib_hip:")


@pytest.mark.gpu
def test_every_method_on_gpu(range_libc):
    """The module's casting methods against the oracle: ray marching (CPU and
    GPU classes), Bresenham, pruned CDDT with the radial pair queries, and
    GiantLUTCast; batch, repeated-angle, sensor-model and fused calls."""
    img = W.corridor_map(300, seed=4)
    omap = range_libc.PyOMap(img)
    R, td = 200.0, 108
    ins = W.world_poses(img, 400, seed=3)
    ang = W.beam_angles(30)
    table = W.sensor_table(201)
    cases = (("PyRayMarchingGPU", "rm", ()), ("PyRayMarching", "rm", ()), ("PyBresenhamsLine", "bl", ()),
             ("PyCDDTCast", "cddt", (td,)), ("PyGiantLUTCast", "glt", (td,)))
    for cls, kind, extra in cases:
        meth = getattr(range_libc, cls)(omap, R, *extra)
        if kind == "cddt":
            meth.prune()
        orc = O.Oracle(img, R, kind, theta_disc=td if extra else 0, prune=kind == "cddt")
        orc.set_sensor_model(table)
        meth.set_sensor_model(table)
        out = np.empty(400, np.float32)
        meth.calc_range_many(ins, out)
        assert_bits_equal(out, orc.numpy_calc_range(ins), f"{cls} calc_range_many")
        rr = np.empty(400 * 30, np.float32)
        meth.calc_range_repeat_angles(ins, ang, rr)
        want_r = orc.numpy_calc_range_angles(ins, ang)
        assert_bits_equal(rr, want_r, f"{cls} repeat_angles")
        obs = W.noisy_observation(want_r[:30], seed=5, max_range=R)
        w = np.empty(400, np.float64)
        meth.eval_sensor_model(obs, rr, w, 30, 400)
        assert np.array_equal(w, orc.eval_sensor_model(obs, want_r, 30, 400)), cls
        wf = np.empty(400, np.float64)
        meth.calc_range_repeat_angles_eval_sensor_model(ins, ang, obs, wf)
        assert np.array_equal(wf, orc.calc_range_repeat_angles_eval_sensor_model(ins, ang, obs)), cls
        if hasattr(meth, "calc_range"):
            x, y, h = (float(v) for v in ins[7])
            assert np.float32(meth.calc_range(x, y, h)).view(np.uint32) == \
                np.float32(orc.calc_range(x, y, h)).view(np.uint32), cls
        if kind == "cddt":
            rad = np.full(400 * 16, -7.0, np.float32)
            meth.calc_range_many_radial_optimized(16, -1.5, 1.5, ins, rad)
            assert_bits_equal(rad, orc.calc_range_many_radial_optimized(ins, 16, -1.5, 1.5), "radial")
