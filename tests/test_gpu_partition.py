"""partition_minimizers (src/minimizer.rs:116-184) on the device with a
caller's histogram: bucket->bin table and bin count equal the oracle's for
the closed-form cases (uniform, all zero, total below the bin count, one
heavy bucket, one bin) and for random skewed histograms whose running
minimum crosses the device's 1024-bucket blocks."""
import numpy as np
import pytest

import cf1g
from gpu_cases import random_genome
from test_oracle_partition import part

pytestmark = pytest.mark.gpu


def cases():
    out = [(np.ones(65536), 128), (np.zeros(65536), 8), (np.arange(65536), 1)]
    h = np.zeros(65536)
    h[10] = 1000
    out.append((h, 4))
    h = np.zeros(65536)
    h[3] = 5
    out.append((h, 4096))  # total below the bin count: target 0
    r = np.random.default_rng(11)
    for nb in (2, 16, 128, 512, 4096):
        h = r.integers(0, 50, 65536) * (r.random(65536) < 0.3)
        out.append((h, nb))
        h = (r.pareto(1.2, 65536) * 100).astype(np.int64)  # heavy tail
        h[r.integers(0, 65536, 5)] = 10 ** 9  # a few huge buckets
        out.append((h, nb))
    h = np.zeros(65536, np.int64)
    h[1023], h[1024], h[2047] = 7, 7 * 10 ** 6, 3  # mass at block edges
    out.append((h, 64))
    return out


def test_partition_matches_oracle():
    g = cf1g.Cf1g(k=31, num_bins=128)
    g.load_records(random_genome(seed=77, n_rec=3, max_len=4000))
    g.count_minimizer_histogram(want=False)
    for i, (h, nb) in enumerate(cases()):
        want_b2b, want_n = part(h, nb)
        b2b, n = g.partition_minimizers_and_route(hist=h, num_bins=nb)
        assert n == want_n, f"case {i}: bins"
        np.testing.assert_array_equal(b2b, want_b2b, err_msg=f"case {i}")
    g.close()
