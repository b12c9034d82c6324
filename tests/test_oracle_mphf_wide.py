"""Wide keys (k > 32) in the project's BBHash: a pair of canonical 63-mers
from the genome-like workload whose linear fold lo ^ hi * C is equal (they
differ at bases 0 and 32). With one level hash state h1 per key, equal h1
would put both keys on the same bit at every level, so the build could never
place them; h1 mixes the high word first, which separates them."""
import numpy as np

from oracle import RefPipeline

PAIR = [b"TGAAAATTGGGGACTCGCTGGTTCCGAGTGGCACATAGTAAGGAGTCGTTGCACCTGGAAACA",
        b"AGAAAATTGGGGACTCGCTGGTTCCGAGTGGCTCATAGTAAGGAGTCGTTGCACCTGGAAACA"]
M = (1 << 64) - 1


def words(s):
    v = 0
    for ch in s:
        v = (v << 2) | "ACGT".index(chr(ch))
    return v & M, v >> 64


def mix64(z):
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M
    return z ^ (z >> 31)


def test_pair_has_equal_linear_fold():
    (l0, h0), (l1, h1) = words(PAIR[0]), words(PAIR[1])
    assert (l0, h0) != (l1, h1)
    assert l0 ^ ((h0 * 0xC2B2AE3D27D4EB4F) & M) == l1 ^ ((h1 * 0xC2B2AE3D27D4EB4F) & M)
    assert l0 ^ mix64(h0) != l1 ^ mix64(h1)


def test_pair_in_one_segment_builds():
    # one bin, one segment: both keys must be placed by the same BBHash
    ref = RefPipeline(63, num_bins=1, threads=2, target_seg=1 << 24)
    ref.load(PAIR)
    assert ref.run_p1_p4() == 2
    lo, hi = ref.keys()
    v = ref.mphf_lookup(lo, hi)
    assert sorted(v.tolist()) == [0, 1]
    assert np.count_nonzero(ref.states()) == 1
