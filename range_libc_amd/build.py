"""Build the HIP library and its data files in-tree.

    python -m range_libc_amd.build            # library + trig table
    python -m range_libc_amd.build --force    # regenerate the trig table too

Outputs (git-ignored, shipped with the source tree):
  range_libc_amd/librangelib_hip.so        the C-ABI library (gfx950)
  range_libc_amd/_data/trig_exc.bin        sinf/cosf exception table
  range_libc_amd/pywrapper/_range_libc.*.so the Cython module over the C ABI
The trig table depends only on rl_trig.h, its generator and the host C
library, so it is regenerated only when one of those changes.
"""
from __future__ import annotations

import hashlib
import os
import platform
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(PKG, "csrc")
DATA = os.path.join(PKG, "_data")
LIB = os.path.join(PKG, "librangelib_hip.so")
TABLE = os.path.join(DATA, "trig_exc.bin")
ROOT = os.path.dirname(PKG)

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("RL_OFFLOAD_ARCH", "gfx950")

# Parity-critical flags: no contraction, no fast-math, correctly rounded
# single-precision divide/sqrt.  Packed FP32 instructions are disabled.
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off", "-fno-fast-math", "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument", "-Wno-unknown-pragmas",
]


def _run(cmd, **kw):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True, **kw)


def _hash_files(paths, extra=""):
    h = hashlib.sha256(extra.encode())
    for p in paths:
        with open(p, "rb") as f:
            h.update(f.read())
    return h.hexdigest()


def build_trig_table(force: bool = False) -> str:
    os.makedirs(DATA, exist_ok=True)
    srcs = [os.path.join(CSRC, "rl_trig.h"), os.path.join(CSRC, "gen_trig_table.cpp")]
    libc = " ".join(platform.libc_ver())
    stamp = _hash_files(srcs, libc)
    stamp_path = TABLE + ".stamp"
    if not force and os.path.exists(TABLE) and os.path.exists(stamp_path):
        with open(stamp_path) as f:
            if f.read().strip() == stamp:
                print(f"trig table up to date ({TABLE})")
                return TABLE
    gen = os.path.join(DATA, "gen_trig_table")
    _run(["g++", "-O2", "-std=c++17", "-fopenmp", "-ffp-contract=off", "-fno-fast-math",
          "-Wno-unknown-pragmas", "-I", CSRC, srcs[1], "-o", gen, "-lm"])
    tmp = TABLE + ".tmp"
    _run([gen, tmp])
    os.replace(tmp, TABLE)
    with open(stamp_path, "w") as f:
        f.write(stamp + "\n")
    return TABLE


def build_library() -> str:
    src = os.path.join(CSRC, "rl_lib.hip")
    tmp = LIB + ".tmp"
    _run([HIPCC, *HIP_FLAGS, "-I", os.path.join(ROOT, "include"), "-I", CSRC, src, "-o", tmp, "-lz"])
    os.replace(tmp, LIB)
    return LIB


def build_cython() -> str:
    """The Cython module over the C ABI (pywrapper/RangeLibc.pyx), linked
    against the library in this directory.  It is built as `_range_libc`;
    `import range_libc` goes through pywrapper/range_libc.py, which lets torch
    (when installed) load its HIP runtime first, as _lib.py does, so that the
    library binds to that runtime instead of bringing in a second one."""
    import sysconfig

    import numpy as np

    pyx = os.path.join(PKG, "pywrapper", "RangeLibc.pyx")
    build_dir = os.path.join(PKG, "_build")
    os.makedirs(build_dir, exist_ok=True)
    c_file = os.path.join(build_dir, "_range_libc.c")
    out = os.path.join(PKG, "pywrapper", "_range_libc" + sysconfig.get_config_var("EXT_SUFFIX"))
    _run([sys.executable, "-m", "cython", "-3", "--module-name", "_range_libc", "-o", c_file, pyx])
    tmp = out + ".tmp"
    _run(["gcc", "-O2", "-shared", "-fPIC", "-w", "-DNPY_NO_DEPRECATED_API=NPY_1_7_API_VERSION",
          "-I", sysconfig.get_paths()["include"], "-I", np.get_include(), "-I", os.path.join(ROOT, "include"),
          c_file, "-o", tmp, "-L", PKG, "-lrangelib_hip", "-Wl,-rpath,$ORIGIN/.."])
    os.replace(tmp, out)
    return out


def build(force_table: bool = False):
    if shutil.which(HIPCC) is None and not os.path.exists(HIPCC):
        raise RuntimeError(f"hipcc not found ({HIPCC})")
    build_trig_table(force_table)
    build_library()
    build_cython()


if __name__ == "__main__":
    build(force_table="--force" in sys.argv)
