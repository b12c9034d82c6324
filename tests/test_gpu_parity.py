"""GPU parity: every phase product of the HIP path equals the CPU oracle's.

All calls go through the C ABI (libcf1g.so via ctypes). Compared exactly:
P1 histogram, P2 bucket->bin table and per-bin routed k-mer multisets, P3
consolidated key array (bins in order, each sorted) and the serialised BBHash,
P4 packed state words, and the P5 outputs (cf_seg / cf_seq line sets, JSON).
"""
import json
import os
import random
import tempfile

import numpy as np
import pytest

import cf1g
from cf1g import synth
from gpu_cases import edge_records, random_genome
from oracle import RefPipeline

pytestmark = pytest.mark.gpu


def decode_gpu_bins(rw, bin_start, recs):
    """Super k-mer strings of the device bin records (flanking bases dropped)."""
    out = []
    for b in range(len(bin_start) - 1):
        strs = []
        for r in range(int(bin_start[b]), int(bin_start[b + 1])):
            w = recs[r]
            L = int(w[-1]) & 255
            v = 0
            for x in w:
                v = (v << 64) | int(x)
            tot = 64 * rw
            # position 0 is the left flanking base; the super k-mer is 1..L
            strs.append("".join("ACGT"[(v >> (tot - 2 * (t + 1))) & 3] for t in range(1, L + 1)))
        out.append(strs)
    return out


def decode_ref_bin(data):
    strs = []
    p = 0
    while p + 2 <= len(data):
        L = data[p] | (data[p + 1] << 8)
        p += 2
        nb = (L + 3) // 4
        strs.append("".join("ACGT"[(data[p + t // 4] >> (6 - 2 * (t % 4))) & 3] for t in range(L)))
        p += nb
    return strs


def kmers_of(strs, k):
    out = []
    for s in strs:
        out.extend(s[i:i + k] for i in range(len(s) - k + 1))
    return sorted(out)


def run_pair(seqs, k, num_bins=128, names=None, p5=True, bins=True, threads=4, target_seg=0, on_done=None):
    names = names or [f"r{i}" for i in range(len(seqs))]
    ref = RefPipeline(k, num_bins=num_bins, threads=threads, target_seg=target_seg)
    ref.load(seqs, names)
    h_ref = ref.p1()
    b2b_ref, nb_ref = ref.p2()
    d_ref = ref.p3()
    ref.p4()

    g = cf1g.Cf1g(k=k, num_bins=num_bins, p3_target_seg=target_seg)
    g.load_records(seqs)
    h = g.count_minimizer_histogram()
    np.testing.assert_array_equal(h, h_ref, err_msg="P1 histogram")
    b2b, nb = g.partition_minimizers_and_route()
    np.testing.assert_array_equal(b2b, b2b_ref, err_msg="P2 bucket->bin")
    assert nb == nb_ref
    if bins:
        rw, bstart, recs = g.bins_export()
        gb = decode_gpu_bins(rw, bstart, recs)
        assert len(gb) == nb_ref
        for b in range(nb_ref):
            # Record boundaries may differ only if chunking differed; the k-mer
            # multiset per bin is what P3 consumes.
            same = kmers_of(gb[b], k) == kmers_of(decode_ref_bin(ref.bin_bytes(b)), k)
            assert same, f"bin {b}: routed k-mers differ"
    d = g.build_mphf()
    assert d == d_ref, "D"
    lo, hi = g.keys()
    rlo, rhi = ref.keys()
    np.testing.assert_array_equal(lo, rlo, err_msg="keys (lo)")
    if k > 32:
        np.testing.assert_array_equal(hi, rhi, err_msg="keys (hi)")
    np.testing.assert_array_equal(g.mphf_export(), ref.mphf_export(), err_msg="BBHash structure")
    g.classify_vertices()
    np.testing.assert_array_equal(g.states(), ref.states(), err_msg="P4 state words")
    if p5:
        with tempfile.TemporaryDirectory() as td:
            pg, pr = os.path.join(td, "gpu"), os.path.join(td, "ref")
            mg = g.p5_write(pg, names=names, input_desc="in.fa", threads=threads)
            ref.p5(pr, input_desc="in.fa")
            for ext in (".cf_seg", ".cf_seq"):
                a = sorted(open(pg + ext).read().splitlines())
                b = sorted(open(pr + ext).read().splitlines())
                assert a == b, ext
            jg = json.load(open(pg + ".json"))
            jr = json.load(open(pr + ".json"))
            jg["parameters info"].pop("output prefix")
            jr["parameters info"].pop("output prefix")
            assert jg == jr
            assert mg == ref.meta()
    if on_done:
        on_done(g, ref)
    g.close()
    return d


@pytest.mark.parametrize("k", [1, 3, 5, 7, 9, 13, 15, 17, 21, 31, 33, 45, 63])
def test_random_genome_all_phases(k):
    run_pair(random_genome(seed=100 + k, n_rec=5, max_len=3000), k, num_bins=16)


def test_edge_records_k31():
    run_pair(edge_records(), 31, num_bins=8)


@pytest.mark.parametrize("k", [5, 21, 63])
def test_edge_records_other_k(k):
    run_pair(edge_records(), k, num_bins=8)


@pytest.mark.parametrize("num_bins", [1, 2, 7, 512, 4096])
def test_bin_counts(num_bins):
    run_pair(random_genome(seed=5, n_rec=3, max_len=20000), 31, num_bins=num_bins, p5=False)


def test_empty_input():
    run_pair([], 31, num_bins=4)


def test_all_short_records():
    run_pair([b"ACGT", b"", b"GGGGGGGGG"], 31, num_bins=4)


def test_heavy_duplicates_bigsort():
    # one 31-mer (poly-A) and four period-4 31-mers repeated ~10^5 times: P3
    # segments exceed the LDS capacity and take the global-memory sort.
    seqs = [b"A" * 120000, b"ACGT" * 30000, b"C" * 5000 + b"N" + b"G" * 5000]
    run_pair(seqs, 31, num_bins=4, bins=False)
    run_pair(seqs, 63, num_bins=4, bins=False)


@pytest.mark.parametrize("k", [31, 63])
def test_segments_above_first_dedup_capacity(k):
    # 2 bins and a 4,000 target give 8 segments of ~1,800-2,900 distinct keys
    # (~1.17 occurrences each): most exceed the 256-thread dedup's 2,048
    # occurrences and take the 512-thread list dedup (up to 4,096), so none
    # is sorted through global memory
    seqs = random_genome(seed=78, n_rec=4, max_len=20000)
    got = {}

    def check(g, ref):
        w = ref.mphf_export()
        nseg, off = int(w[5]), 16 + 65536 // 4
        got["u"] = np.diff(w[off:off + nseg + 1].astype(np.int64))
        got["info"] = g.info()

    run_pair(seqs, k, num_bins=2, target_seg=4000, p5=False, bins=False, on_done=check)
    assert (got["u"] > 2048).sum() >= 6 and got["u"].max() < 3000
    assert got["info"]["n_oversized"] == 0


def test_transcriptome_small():
    a, off, names = synth.transcriptome(scale=0.003, seed=2, k=31)
    seqs = synth.records_list(a, off)
    run_pair(seqs, 31, num_bins=64, names=names, bins=False)
    run_pair(seqs, 21, num_bins=64, names=names, bins=False)


def test_genome_5m_scaled():
    a, off, names = synth.genome_5m(scale=0.05)
    run_pair(synth.records_list(a, off), 31, num_bins=128, names=names, bins=False)


def test_genome_300m_shape_scaled():
    a, off, names = synth.genome_300m(scale=0.001)
    run_pair(synth.records_list(a, off), 31, num_bins=32, names=names, bins=False)


@pytest.mark.parametrize("k", [15, 31, 63])
def test_mphf_segment_sizes(k):
    seqs = random_genome(seed=77, n_rec=4, max_len=30000)
    # one segment per bin, > register capacity: global-memory MPHF build
    run_pair(seqs, k, num_bins=2, target_seg=1 << 24, p5=False, bins=False)
    # 2^12 segments per bin, mostly tiny
    run_pair(seqs, k, num_bins=8, target_seg=4, p5=False, bins=False)


def test_mphf_lookup_batch():
    seqs = random_genome(seed=9, n_rec=3, max_len=8000)
    for k in (31, 63):
        g = cf1g.Cf1g(k=k, num_bins=8)
        g.load_records(seqs)
        d = g.run_p1_p4()
        lo, hi = g.keys()
        v = g.mphf_lookup(lo, hi)
        assert v.tolist() == list(range(d)), "keys are stored in value order; BBHash is a bijection onto [0, D)"
        g.close()


def test_phase_order_is_enforced():
    g = cf1g.Cf1g(k=31, num_bins=8)
    with pytest.raises(cf1g.Cf1gError) as e:
        g.build_mphf()
    assert e.value.code == -2
    g.close()


def test_wide_keys_with_equal_linear_fold():
    # two 63-mers of the genome-like workload whose linear fold lo ^ hi * C is
    # equal, in one segment: the BBHash must still place both
    from test_oracle_mphf_wide import PAIR
    run_pair(PAIR, 63, num_bins=1, target_seg=1 << 24, bins=False)
    run_pair(PAIR + [PAIR[0] + b"ACGTTGCA" + PAIR[1]], 63, num_bins=2, bins=False)


def graph_by_kmer(g, k):
    """States keyed by canonical k-mer and the canonicalised P5 output
    (cf_seg strings; cf_seq tilings with ids replaced by those strings):
    what does not depend on the minimizer scheme, bins or MPHF."""
    lo, hi = g.keys()
    words = g.states()
    v = np.arange(len(lo), dtype=np.uint64)
    st = (words[v // 12] >> ((v % 12) * 5).astype(np.uint64)) & np.uint64(31)
    keys = list(zip(hi.tolist(), lo.tolist())) if k > 32 else lo.tolist()
    states = dict(zip(keys, st.tolist()))
    with tempfile.TemporaryDirectory() as td:
        p = os.path.join(td, "o")
        g.p5_write(p)
        seg = dict(line.split("\t") for line in open(p + ".cf_seg").read().splitlines())
        tiles = []
        for line in open(p + ".cf_seq").read().splitlines():
            name, toks = line.split("\t")
            tiles.append((name, [t if t.startswith("N") else (seg[t[:-1]], t[-1]) for t in toks.split()]))
    return states, sorted(seg.values()), tiles


@pytest.mark.parametrize("k,m", [(31, 5), (21, 3), (63, 5), (63, 9)])
def test_minimizer_length_invariance(k, m):
    # m = 5 at k = 31 gives super k-mers of up to 57 bases, the record capacity
    seqs = random_genome(seed=60 + k + m, n_rec=4, max_len=6000)
    out = []
    for mm in (0, m):
        g = cf1g.Cf1g(k=k, num_bins=8, m=mm)
        g.load_records(seqs)
        g.run_p1_p4()
        out.append(graph_by_kmer(g, k))
        g.close()
    assert out[0] == out[1]


@pytest.mark.parametrize("k", [15, 31, 63])
def test_mphf_lookup_absent_keys(k):
    # keys outside the set: the device's values equal the oracle's (0 when a
    # key misses every level of its segment; levels are read only inside the
    # segment's blocks), for many segments and for one large one
    seqs = random_genome(seed=31 + k, n_rec=3, max_len=6000)
    rng = np.random.default_rng(k)
    for bins, tgt in ((8, 0), (1, 1 << 24)):
        ref = RefPipeline(k, num_bins=bins, threads=2, target_seg=tgt)
        ref.load(seqs)
        ref.run_p1_p4()
        g = cf1g.Cf1g(k=k, num_bins=bins, p3_target_seg=tgt)
        g.load_records(seqs)
        g.run_p1_p4()
        mask = (1 << (2 * k)) - 1 if k <= 32 else (1 << 64) - 1
        lo = rng.integers(0, 1 << 63, 20000, dtype=np.uint64) & np.uint64(mask)
        hi = (rng.integers(0, 1 << 62, 20000, dtype=np.uint64) & np.uint64((1 << (2 * k - 64)) - 1)) if k > 32 else None
        got = g.mphf_lookup(lo, hi)
        np.testing.assert_array_equal(got, ref.mphf_lookup(lo, hi))
        assert got.max() < g.info()["n_distinct"]
        g.close()
