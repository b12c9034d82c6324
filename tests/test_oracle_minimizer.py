"""The project's minimizer scheme in the oracle (P1 histogram, P2 routing)
against a brute-force restatement of its definition, plus the property per-bin
dedup relies on: a k-mer and its reverse complement route to the same bin
(src/mphf.rs:27-28)."""
import random

import numpy as np
import pytest

from conftest import rand_seq, rc
from oracle import RefPipeline, lib

ACGT = "ACGT"


def minimizer_len(k):
    m = min(k // 2, 15)
    return max(m - 1, 1) if m % 2 == 0 else m


def enc(s):
    v = 0
    for c in s:
        v = (v << 2) | ACGT.index(c.upper())
    return v


def canon_val(s):
    return min(enc(s), enc(rc(s)))


def valid(s, i, k):
    return i >= 0 and i + k <= len(s) and all(c in "ACGTacgt" for c in s[i:i + k])


def min_pos(s, i, k, m):
    w = k - m + 1
    best = None
    for j in range(i, i + w):
        o = lib().ref_mmer_order(canon_val(s[j:j + m]))
        if best is None or o < best[0]:
            best = (o, j)
    return best[1]


def brute_superkmers(s, k):
    m = minimizer_len(k)
    out = []
    prev = None
    for i in range(len(s) - k + 1):
        if not valid(s, i, k):
            prev = None
            continue
        p = min_pos(s, i, k, m)
        if prev is None or p != prev:
            out.append([i, 1, canon_val(s[p:p + m])])
        else:
            out[-1][1] += 1
        prev = p
    return out


@pytest.mark.parametrize("k", [1, 3, 5, 9, 15, 21, 31, 33, 63])
def test_histogram_matches_definition(k):
    r = random.Random(k)
    seqs = [rand_seq(r, r.randint(0, 400), pn=0.02, plow=0.1) for _ in range(4)]
    hist = np.zeros(65536, np.uint64)
    for s in seqs:
        for _, _, v in brute_superkmers(s, k):
            hist[lib().ref_bucket_hash(v)] += 1
    p = RefPipeline(k, num_bins=8)
    p.load([s.encode() for s in seqs])
    np.testing.assert_array_equal(p.p1(), hist)


def decode_bin(data):
    strs, p = [], 0
    while p + 2 <= len(data):
        L = data[p] | (data[p + 1] << 8)
        p += 2
        strs.append("".join(ACGT[(data[p + t // 4] >> (6 - 2 * (t % 4))) & 3] for t in range(L)))
        p += (L + 3) // 4
    return strs


@pytest.mark.parametrize("k", [5, 15, 31, 63])
def test_routing_matches_definition(k):
    r = random.Random(50 + k)
    seqs = [rand_seq(r, r.randint(100, 600), pn=0.01) for _ in range(3)]
    p = RefPipeline(k, num_bins=16)
    p.load([s.encode() for s in seqs])
    p.p1()
    b2b, nb = p.p2()
    want = {}
    for s in seqs:
        for i, n, v in brute_superkmers(s, k):
            want.setdefault(int(b2b[lib().ref_bucket_hash(v)]), []).append(s[i:i + n - 1 + k].upper())
    for b in range(nb):
        assert sorted(decode_bin(p.bin_bytes(b))) == sorted(want.get(b, [])), b


@pytest.mark.parametrize("k", [3, 15, 31, 45, 63])
def test_strand_symmetric_minimizer_value(k):
    r = random.Random(k * 7)
    m = minimizer_len(k)
    for _ in range(200):
        x = rand_seq(r, k)
        y = rc(x)
        vx = canon_val(x[min_pos(x, 0, k, m):][:m])
        vy = canon_val(y[min_pos(y, 0, k, m):][:m])
        assert vx == vy
