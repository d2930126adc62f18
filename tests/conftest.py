import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the C ABI on the device)")
    config.addinivalue_line("markers", "slow: long-running CPU check")


@pytest.hookimpl(trylast=True)
def pytest_collection_modifyitems(config, items):
    """GPU tests hold inputs in torch tensors, and torch ships its own HIP
    runtime beside the system one the library links.  torch's runtime has
    been seen to find no device once the library's had opened it, while the
    other order works, so torch initialises the device before any test
    creates a context."""
    if any(item.get_closest_marker("gpu") for item in items):
        import torch

        torch.cuda.is_available()


def bits(a):
    """uint32 view for bit-exact float comparisons (NaN-safe)."""
    return np.ascontiguousarray(a, dtype=np.float32).view(np.uint32)


def assert_bits_equal(got, want, what=""):
    g, w = bits(got), bits(want)
    assert g.shape == w.shape, f"{what}: shape {g.shape} vs {w.shape}"
    bad = np.nonzero(g != w)[0]
    if bad.size:
        i = bad[0]
        raise AssertionError(f"{what}: {bad.size} of {g.size} values differ; first at {i}: "
                             f"{got.reshape(-1)[i]!r} vs {want.reshape(-1)[i]!r}")


@pytest.fixture(scope="session")
def ctx():
    import range_libc_amd as rl

    return rl.default_context(0)
