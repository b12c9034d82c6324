"""cf1g_mphf_lookup_noalloc: the lookup that allocates nothing (keys and
values staged through a fixed device array) gives exactly cf1g_mphf_lookup's
values, on consecutive calls with shrinking inputs, and refuses more keys
than its array holds."""
import ctypes as C

import numpy as np
import pytest

import cf1g
from gpu_cases import random_genome

pytestmark = pytest.mark.gpu

U64P = C.POINTER(C.c_uint64)


def noalloc(g, keys, n):
    out = np.zeros(max(n, 1), np.uint64)
    rc = g.L.cf1g_mphf_lookup_noalloc(g.h, keys.ctypes.data_as(U64P), n, out.ctypes.data_as(U64P))
    return rc, out[:n]


@pytest.mark.parametrize("k", [31, 63])
def test_noalloc_lookup_equals_lookup(k):
    seqs = random_genome(seed=70 + k, n_rec=4, max_len=40000)
    g = cf1g.Cf1g(k=k, num_bins=16)
    g.load_records(seqs)
    d = g.run_p1_p4()
    lo, hi = g.keys()
    rng = np.random.default_rng(k)
    words = 2 if k > 32 else 1
    for n in (d, d // 3 + 1, 1):  # the second call is smaller, then a single key
        idx = rng.integers(0, d, n)
        plo = lo[idx]
        phi = hi[idx] if k > 32 else None
        want = g.mphf_lookup(plo, phi)
        keys = np.empty(words * n, np.uint64)
        keys[0::words] = plo
        if k > 32:
            keys[1::2] = phi
        rc, got = noalloc(g, keys, n)
        assert rc == 0, g.L.cf1g_last_error(g.h).decode()
        np.testing.assert_array_equal(got, want)
        np.testing.assert_array_equal(got, idx.astype(np.uint64))  # keys are in value order
    # absent keys: the same values as the allocating lookup
    alo = rng.integers(0, 1 << 62, 5000, dtype=np.uint64)
    ahi = rng.integers(0, 1 << 60, 5000, dtype=np.uint64) if k > 32 else None
    keys = np.empty(words * 5000, np.uint64)
    keys[0::words] = alo
    if k > 32:
        keys[1::2] = ahi
    rc, got = noalloc(g, keys, 5000)
    assert rc == 0
    np.testing.assert_array_equal(got, g.mphf_lookup(alo, ahi))
    # beyond the fixed array: refused with a clear error, nothing written
    cap = (32 << 20) // (8 * words + 8) - 64
    big = np.zeros(words * (cap + 1), np.uint64)
    rc, _ = noalloc(g, big, cap + 1)
    assert rc == -1  # CF1G_E_INVALID
    assert "at most" in g.L.cf1g_last_error(g.h).decode()
    g.close()


def test_noalloc_lookup_needs_mphf():
    g = cf1g.Cf1g(k=31, num_bins=4)
    keys = np.zeros(4, np.uint64)
    rc, _ = noalloc(g, keys, 4)
    assert rc == -2  # CF1G_E_ORDER: no MPHF yet
    g.close()
