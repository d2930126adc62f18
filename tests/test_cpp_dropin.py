"""The C++ drop-in header compiles and links like reference code using RangeLib.h."""
import os
import subprocess

import pytest

from range_libc_amd import _lib
from tests.conftest import ROOT

SRC = os.path.join(ROOT, "tests", "cpp", "drop_in_example.cpp")


def _build(tmp_path):
    exe = str(tmp_path / "drop_in_example")
    cmd = ["g++", "-std=c++11", "-O1", "-I", os.path.join(ROOT, "include"), SRC, "-o", exe,
           "-L", _lib.PKG, "-lrangelib_hip", f"-Wl,-rpath,{_lib.PKG}"]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    return exe


def test_header_compiles_and_links(tmp_path):
    assert os.path.exists(_build(tmp_path))


def test_no_device_raises_runtime_error(tmp_path):
    exe = _build(tmp_path)
    env = dict(os.environ, HIP_VISIBLE_DEVICES="-1", ROCR_VISIBLE_DEVICES="-1")
    p = subprocess.run([exe], capture_output=True, text=True, env=env, timeout=120)
    assert p.returncode == 3, p.stdout + p.stderr
    assert "runtime_error: rangelib_hip:" in p.stdout


@pytest.mark.gpu
def test_drop_in_on_gpu(tmp_path):
    exe = _build(tmp_path)
    p = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stdout + p.stderr
    # wall at y = 15 (values from the oracle): from (3.5, 2.5) heading fl(pi/2) the hit
    # cell is (3, 15); from (8, 8) the tiny negative cosine truncates x to 7, cell (7, 15)
    assert p.stdout.split() == ["rm", "12.509996", "7.071068", "bl", "13.000000"]


PNG_SRC = os.path.join(ROOT, "tests", "cpp", "png_dropin.cpp")
MAPS = os.path.join(ROOT, "tests", "golden", "maps")


def _build_png(tmp_path):
    exe = str(tmp_path / "png_dropin")
    cmd = ["g++", "-std=c++11", "-O1", "-I", os.path.join(ROOT, "include"), PNG_SRC, "-o", exe,
           "-L", _lib.PKG, "-lrangelib_hip", f"-Wl,-rpath,{_lib.PKG}"]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    return exe


@pytest.mark.parametrize("name,key", [("basement_hallways_5cm.png", "basement_hallways_5cm"),
                                      ("synthetic.map.png", "synthetic_map"), ("small.map.png", "small_map")])
def test_cpp_omap_from_png(tmp_path, name, key):
    """ranges::OMap(filename, threshold) (RangeLib.h:158-201) reads the
    reference's maps as the compiled reference does (golden occupancy)."""
    import numpy as np

    exe = _build_png(tmp_path)
    z = np.load(os.path.join(ROOT, "tests", "golden", "png_maps.npz"))
    for thr in (128, 1):
        shape = tuple(int(v) for v in z[f"{key}_t{thr}_shape"])
        want = np.unpackbits(z[f"{key}_t{thr}_occ"])[: shape[0] * shape[1]].reshape(shape)
        out = subprocess.run([exe, os.path.join(MAPS, name), str(thr), "info"], capture_output=True, text=True,
                             check=True, timeout=60).stdout.split()
        assert [int(v) for v in out] == [want.shape[1], want.shape[0], 0, int(want.sum())]


def test_cpp_omap_png_error(tmp_path):
    exe = _build_png(tmp_path)
    out = subprocess.run([exe, str(tmp_path / "missing.png"), "128", "info"], capture_output=True, text=True,
                         check=True, timeout=60).stdout
    assert "ERROR" in out and out.split()[-4:] == ["1", "1", "1", "0"]


@pytest.mark.gpu
def test_cpp_png_map_methods_on_gpu(tmp_path):
    """A PNG-loaded ranges::OMap through RayMarching, BresenhamsLine, pruned
    CDDTCast and GiantLUTCast (numpy_calc_range), bit for bit against the
    oracle on the reference's basement map."""
    import numpy as np

    from oracle import oracle as O
    from range_libc_amd import workloads as W

    exe = _build_png(tmp_path)
    z = np.load(os.path.join(ROOT, "tests", "golden", "png_maps.npz"))
    shape = tuple(int(v) for v in z["basement_hallways_5cm_t128_shape"])
    img = np.unpackbits(z["basement_hallways_5cm_t128_occ"])[: shape[0] * shape[1]].reshape(shape)
    q = W.world_poses(img, 3000, seed=17)
    R, td = 300.0, 108
    text = "\n".join(" ".join(repr(float(v)) for v in row) for row in q.astype(np.float64))
    p = subprocess.run([exe, os.path.join(MAPS, "basement_hallways_5cm.png"), "128", "cast", str(R), str(td)],
                       input=text, capture_output=True, text=True, timeout=300)
    assert p.returncode == 0, p.stdout[-2000:] + p.stderr[-2000:]
    got = np.array([[int(v, 16) for v in line.split()] for line in p.stdout.split("\n") if line.strip()],
                   dtype=np.uint64).astype(np.uint32)
    assert got.shape == (3000, 4)
    wants = [O.Oracle(img, R, "rm").numpy_calc_range(q), O.Oracle(img, R, "bl").numpy_calc_range(q),
             O.Oracle(img, R, "cddt", theta_disc=td, prune=True).numpy_calc_range(q),
             O.Oracle(img, R, "glt", theta_disc=td).numpy_calc_range(q)]
    for k, (name, want) in enumerate(zip(("rm", "bl", "pcddt", "glt"), wants)):
        bad = np.nonzero(got[:, k] != want.view(np.uint32))[0]
        assert bad.size == 0, f"{name}: {bad.size} differ, first {bad[:5]}"
