"""The CDDT build's device-wide exclusive sum and per-segment float sort
(hand-written kernels) against numpy, over every size class of the sort:
one-wave segments (<= 256 keys), 256-thread (<= 2048) and 1024-thread
(<= 16384) workgroups, and the global-memory bitonic path above that; the
scan over one tile, several tiles and more tiles than its spine's chunk."""
import numpy as np
import pytest

from tests.conftest import assert_bits_equal

pytestmark = pytest.mark.gpu


def _torch():
    import torch

    assert torch.cuda.is_available(), "GPU test needs a device"
    return torch


def _sort_ref(vals, start):
    """Each segment in the order of the radix sort of float bit patterns
    (negative values by inverted bits, -0 before +0), as numpy sorts the
    order-preserving unsigned keys."""
    u = vals.view(np.uint32)
    keys = np.where(u & np.uint32(0x80000000), ~u, u | np.uint32(0x80000000))
    out = np.empty_like(vals)
    for s in range(start.size - 1):
        a, b = start[s], start[s + 1]
        out[a:b] = vals[a:b][np.argsort(keys[a:b], kind="stable")]
    return out


@pytest.mark.parametrize("n", [1, 7, 2047, 2048, 2049, 4_200_000, 5_000_003])
def test_exclusive_sum(ctx, n):
    torch = _torch()
    rng = np.random.default_rng(n)
    a = rng.integers(0, 1 << 20, n, dtype=np.uint64)
    a[rng.integers(0, n, max(1, n // 1000))] = np.uint64(1 << 40)  # a few large counts
    d_in = torch.from_numpy(a.view(np.int64)).cuda()
    d_out = torch.full((n,), -1, dtype=torch.int64, device="cuda")
    ctx.exclusive_sum_dev(d_in.data_ptr(), d_out.data_ptr(), n)
    got = d_out.cpu().numpy().view(np.uint64)
    want = np.concatenate([[0], np.cumsum(a)[:-1]]).astype(np.uint64)
    assert np.array_equal(got, want), f"first difference at {np.nonzero(got != want)[0][:1]}"


def test_segmented_sort_size_classes(ctx):
    torch = _torch()
    rng = np.random.default_rng(7)
    sizes = [0, 1, 2, 3, 0, 64, 65, 255, 256, 257, 0, 1000, 2047, 2048, 2049, 5000, 16383, 16384, 16385, 40000, 1, 0]
    sizes += list(rng.integers(0, 400, 3000))  # many short segments, as the CDDT bins of a map are
    start = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    n = int(start[-1])
    vals = (rng.standard_normal(n) * 300).astype(np.float32)
    vals[rng.integers(0, n, n // 50)] = np.float32(0.0)
    vals[rng.integers(0, n, n // 50)] = np.float32(-0.0)
    vals[rng.integers(0, n, n // 100)] = np.float32(np.inf)
    vals[rng.integers(0, n, n // 100)] = np.float32(-np.inf)
    vals[rng.integers(0, n, n // 20)] = np.float32(12.5)  # runs of duplicates
    d_in = torch.from_numpy(vals).cuda()
    d_out = torch.full((n,), float("nan"), dtype=torch.float32, device="cuda")
    d_start = torch.from_numpy(start).cuda()
    ctx.segmented_sort_dev(d_in.data_ptr(), d_out.data_ptr(), d_start.data_ptr(), start.size - 1)
    assert_bits_equal(d_out.cpu().numpy(), _sort_ref(vals, start), "segmented sort")
    assert np.array_equal(d_in.cpu().numpy().view(np.uint32), vals.view(np.uint32)), "input modified"


def test_segmented_sort_empty_and_single(ctx):
    torch = _torch()
    start = np.zeros(5, np.int64)
    d_start = torch.from_numpy(start).cuda()
    d_buf = torch.zeros(1, dtype=torch.float32, device="cuda")
    ctx.segmented_sort_dev(d_buf.data_ptr(), d_buf.data_ptr(), d_start.data_ptr(), 4)  # all segments empty
    ctx.segmented_sort_dev(d_buf.data_ptr(), d_buf.data_ptr(), d_start.data_ptr(), 0)  # no segments
    vals = np.array([3.0, -1.0, 2.0], np.float32)
    start = np.array([0, 1, 3], np.int64)
    d_in = torch.from_numpy(vals).cuda()
    d_out = torch.empty(3, dtype=torch.float32, device="cuda")
    ctx.segmented_sort_dev(d_in.data_ptr(), d_out.data_ptr(), torch.from_numpy(start).cuda().data_ptr(), 2)
    assert_bits_equal(d_out.cpu().numpy(), np.array([3.0, -1.0, 2.0], np.float32), "two short segments")
