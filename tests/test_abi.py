"""CPU checks of the C-ABI library: it loads and exports what the header declares.

No compute calls: loading librangelib_hip.so needs no GPU.
"""
import ctypes as C
import os
import re
import subprocess
import sys

import pytest

import range_libc_amd as rl
from range_libc_amd import _lib
from tests.conftest import ROOT

HEADER = os.path.join(ROOT, "include", "rangelib_hip.h")


def _declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(rl_[a-z0-9_]+)\s*\(", text)))


def test_library_loads():
    assert os.path.exists(_lib.LIB_PATH), "run the build first"
    h = _lib.lib()
    assert h.rl_abi_version() == 1


def test_exports_every_declared_symbol():
    names = _declared()
    assert len(names) >= 25
    h = _lib.lib()
    missing = [n for n in names if not hasattr(h, n)]
    assert not missing, missing


def test_python_binding_covers_header():
    names = set(_declared()) - {"rl_last_error"}
    assert names <= set(_lib.SIGNATURES), sorted(names - set(_lib.SIGNATURES))


def test_null_arguments_rejected_without_device():
    h = _lib.lib()
    # argument checks run before any device call
    assert h.rl_set_stream(None, None) == 1
    assert b"ctx" in h.rl_last_error()
    assert h.rl_map_set_world(None, 1, 0, 0, 0, 0, 1) == 1
    assert h.rl_create(0, None) == 1


def test_trig_table_built():
    path = os.path.join(_lib.PKG, "_data", "trig_exc.bin")
    assert os.path.exists(path)
    with open(path, "rb") as f:
        assert f.read(8) == b"RLTRIG01"


def test_module_flags_mirror_reference():
    # RangeLibc.pyx:94-102 with RangeLib.h:64-71 defaults
    assert rl.USE_CACHED_TRIG is False and rl.USE_ALTERNATE_MOD is True
    assert rl.USE_LRU_CACHE is False and rl.LRU_CACHE_SIZE == 1000000
    assert rl.SHOULD_USE_CUDA is True


def test_trig_table_libc_version_matches():
    h = _lib.lib()
    a, b = C.create_string_buffer(64), C.create_string_buffer(64)
    assert h.rl_trig_table_info(a, b, 64) == 0, h.rl_last_error()
    assert a.value == b.value and a.value


_LIBC_CHECK = r"""
import ctypes as C, sys
sys.path.insert(0, sys.argv[1])
from range_libc_amd import _lib
h = _lib.lib()
a, b = C.create_string_buffer(64), C.create_string_buffer(64)
rc = h.rl_trig_table_info(a, b, 64)
print(rc, a.value.decode(), b.value.decode(), h.rl_last_error().decode(), sep="|")
"""


def test_trig_table_from_another_libc_is_refused(tmp_path):
    """A table generated with another C library release is reported (and
    rl_create refuses it) unless the mismatch is explicitly allowed."""
    src = os.path.join(_lib.PKG, "_data", "trig_exc.bin")
    dst = tmp_path / "trig_exc.bin"
    data = bytearray(open(src, "rb").read())
    # header: magic[8], 4 x uint32, 2 x double, libc_version[32]
    off = 8 + 16 + 16
    data[off:off + 32] = b"1.99".ljust(32, b"\0")
    dst.write_bytes(bytes(data))
    env = dict(os.environ, RL_TRIG_TABLE=str(dst))
    env.pop("RL_TRIG_ALLOW_LIBC_MISMATCH", None)
    out = subprocess.run([sys.executable, "-c", _LIBC_CHECK, ROOT], env=env, capture_output=True, text=True,
                         check=True).stdout.strip().split("|")
    assert out[0] == "3" and out[1] == "1.99" and out[2] != "1.99"
    assert "1.99" in out[3] and "RL_TRIG_ALLOW_LIBC_MISMATCH" in out[3]
    env["RL_TRIG_ALLOW_LIBC_MISMATCH"] = "1"
    out = subprocess.run([sys.executable, "-c", _LIBC_CHECK, ROOT], env=env, capture_output=True, text=True,
                         check=True).stdout.strip().split("|")
    assert out[0] == "0"
