"""partition_minimizers / Partitioning::from_boundaries (src/minimizer.rs:116-184):
closed-form cases and invariants."""
import ctypes as C

import numpy as np
import pytest

from oracle import lib


def part(hist, nb):
    hist = np.ascontiguousarray(hist, np.uint64)
    b2b = np.zeros(65536, np.uint16)
    n = C.c_uint32(0)
    lib().ref_partition(hist.ctypes.data_as(C.POINTER(C.c_uint64)), nb, b2b.ctypes.data_as(C.POINTER(C.c_uint16)), C.byref(n))
    return b2b, n.value


def test_uniform_histogram():
    # every bucket 1, 128 bins: target 512; a boundary after every 512 buckets
    b2b, n = part(np.ones(65536), 128)
    assert n == 128
    assert np.array_equal(b2b, (np.arange(65536) // 512).astype(np.uint16))


def test_zero_histogram_target_zero():
    # total < num_bins: target 0, boundaries at buckets 0..bins-2 (one per bucket)
    b2b, n = part(np.zeros(65536), 8)
    assert n == 8
    assert np.array_equal(b2b, np.minimum(np.arange(65536), 7).astype(np.uint16))


def test_single_heavy_bucket_carries_excess():
    # all mass in bucket 10: target = total/4; bucket 10 places one boundary
    # and the carried excess places one boundary per following bucket.
    h = np.zeros(65536)
    h[10] = 1000
    b2b, n = part(h, 4)
    assert n == 4
    assert b2b[10] == 0 and b2b[11] == 1 and b2b[12] == 2 and b2b[13] == 3 and b2b[65535] == 3


def test_one_bin():
    b2b, n = part(np.arange(65536), 1)
    assert n == 1 and not b2b.any()


@pytest.mark.parametrize("seed", range(6))
def test_invariants_random(seed):
    r = np.random.default_rng(seed)
    h = r.integers(0, 50, 65536) * (r.random(65536) < 0.3)
    nb = int(r.choice([2, 16, 128, 512, 4096]))
    b2b, n = part(h, nb)
    d = np.diff(b2b.astype(np.int64))
    assert b2b[0] == 0 and np.all((d == 0) | (d == 1))  # monotone, contiguous bins
    assert n <= nb and b2b[-1] == n - 1
    # each bin but the last closes once its count reaches the target
    target = int(h.sum()) // nb
    sums = np.bincount(b2b, weights=h, minlength=n)
    assert np.all(sums[:-1] >= 0)
    if target > 0:
        carry = 0
        for b in range(n - 1):
            carry += sums[b]
            assert carry >= target * (b + 1) - 0  # prefix mass reaches b+1 targets
