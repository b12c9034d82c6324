"""The oracle's P1 / P2 chunking follows the reference's definition.

A record of CHUNK_BASES (4,000,000) bases or more is scanned in chunks that
start every CHUNK_BASES bases, each as a sequence of its own: P1's chunks are
CHUNK_BASES long (src/minimizer.rs:57-86), P2's extend K-1 bases further
(src/superkmer.rs:57-88). So P1's histogram of such a record equals the
histogram of its P1 chunks given as separate records, and P2's routed
records equal those of its P2 chunks given as separate records.
"""
import numpy as np
import pytest

from oracle import RefPipeline

CHUNK = 4_000_000


def chunks(s, overlap):
    if len(s) < CHUNK:
        return [s]
    return [s[a:a + CHUNK + overlap] for a in range(0, len(s), CHUNK)]


def routed(ref, rw=2):
    """Routed records as sorted (base count, words) rows; bins left out (the
    partition depends on each input's own histogram)."""
    _, L, w = ref.bin_records(rw)
    o = np.lexsort([w[:, 1], w[:, 0], L])
    return L[o], w[o]


@pytest.mark.parametrize("extra", [0, 1, 29, 30, 31, 400_000, 500_000])
def test_long_record_is_scanned_in_chunks(extra):
    k = 31
    rng = np.random.default_rng(extra)
    s = bytearray(np.frombuffer(b"ACGT", np.uint8)[rng.integers(0, 4, CHUNK + extra)].tobytes())
    if extra == 500_000:
        s[CHUNK - 10:CHUNK + 5] = b"N" * 15  # an N run across the edge
    s[1000:1040] = s[CHUNK - 100:CHUNK - 60]  # a repeat
    s = bytes(s)
    other = b"ACGTTGCA" * 10
    whole = RefPipeline(k, num_bins=16, threads=4)
    whole.load([s, other])
    h = whole.p1()
    whole.p2()
    p1 = RefPipeline(k, num_bins=16, threads=4)
    p1.load(chunks(s, 0) + [other])
    np.testing.assert_array_equal(h, p1.p1(), err_msg="P1 = the histogram of its chunks")
    p2 = RefPipeline(k, num_bins=16, threads=4)
    p2.load(chunks(s, k - 1) + [other])
    p2.p1()
    p2.p2()
    got, want = routed(p2), routed(whole)
    for x, y in zip(got, want):
        np.testing.assert_array_equal(x, y, err_msg="P2 = the super k-mers of its overlapping chunks")
    if len(s) > CHUNK + k and extra != 500_000:
        # the edge splits runs: P2 routes more super k-mers than P1 counts
        assert int(h.sum()) < len(want[0])
