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
