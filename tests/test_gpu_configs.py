"""GPU parity at the benchmark configurations' full sizes.

Each workload runs through the C ABI and through the CPU oracle on all host
cores, and every phase product is compared exactly: the P1 histogram, the
P2 bucket->bin table and every routed record (bin, bases), D, the keys in
MPHF value order, the serialised BBHash, the P4 state words and the P5
outputs (cf_seg / cf_seq line sets, JSON, meta).

  genome_5m      one 5 Mbp record (200 planted 2 kbp repeats), k=31, 128
                 bins: crosses the 4 Mbp chunk edge of P1 / P2
  transcriptome  100k records (2% shorter than k), k=31 and k=21, 128 bins,
                 short records tracked in the JSON
  genome_300m    24 records, 300 Mbp, poly-N runs, k=31 and k=63, bin count
                 and segment target chosen by the autotuner
  learned index  every one of the genome's distinct 31-mers looked up
                 through the learned index: its exact rank; classification
                 and P5 through the index give the BBHash path's graph
"""
import json
import os
import tempfile

import numpy as np
import pytest

import cf1g
from cf1g import synth
from oracle import RefPipeline

pytestmark = pytest.mark.gpu

THREADS = max(1, min(16, len(os.sched_getaffinity(0))))


@pytest.fixture
def progress(request):
    """Progress lines on the terminal while a long case runs (not captured)."""
    capman = request.config.pluginmanager.getplugin("capturemanager")

    def say(msg):
        with capman.global_and_fixture_disabled():
            os.write(2, f"  [{request.node.name}] {msg}\n".encode())

    return say


@pytest.fixture(scope="module")
def genome300m():
    return synth.genome_300m()


@pytest.fixture(scope="module")
def transcriptome():
    return synth.transcriptome(k=31)


def gpu_records(rw, bin_start, recs):
    """Device bin records -> (bin, base count, rw words of the super k-mer's
    bases MSB-first from bit 63 of word 0), flanking bases and the
    length / flag bits dropped: the form the oracle's bin_records gives."""
    recs = np.asarray(recs, np.uint64)
    n = recs.shape[0]
    nb = len(bin_start) - 1
    b = np.repeat(np.arange(nb, dtype=np.uint64), np.diff(bin_start.astype(np.int64)))
    L = recs[:, rw - 1] & np.uint64(255)
    out = np.empty_like(recs)
    two, s62 = np.uint64(2), np.uint64(62)
    for q in range(rw):
        nxt = recs[:, q + 1] >> s62 if q + 1 < rw else np.zeros(n, np.uint64)
        out[:, q] = (recs[:, q] << two) | nxt  # drop the left flank
    for q in range(rw):
        keep = np.clip(2 * L.astype(np.int64) - 64 * q, 0, 64)  # bits of bases < L in word q
        low = np.left_shift(np.uint64(1), (64 - np.clip(keep, 1, 64)).astype(np.uint64)) - np.uint64(1)
        mask = np.where(keep >= 64, ~np.uint64(0), np.where(keep <= 0, np.uint64(0), ~low))
        out[:, q] &= mask
    return b, L, out


def sort_records(b, L, w):
    keys = [w[:, q] for q in range(w.shape[1] - 1, -1, -1)] + [L, b]
    o = np.lexsort(keys)
    return b[o], L[o], w[o]


def compare_p5(g, ref, names, track_short, tag):
    with tempfile.TemporaryDirectory() as td:
        pg, pr = os.path.join(td, "gpu"), os.path.join(td, "ref")
        mg = g.p5_write(pg, names=names, input_desc="in.fa", threads=THREADS, track_short=track_short)
        ref.p5(pr, input_desc="in.fa", track_short=track_short)
        for ext in (".cf_seg", ".cf_seq"):
            with open(pg + ext) as f:
                a = sorted(f.read().splitlines())
            with open(pr + ext) as f:
                b = sorted(f.read().splitlines())
            assert a == b, f"{tag}: {ext} differs"
        with open(pg + ".json") as f:
            jg = json.load(f)
        with open(pr + ".json") as f:
            jr = json.load(f)
        jg["parameters info"].pop("output prefix")
        jr["parameters info"].pop("output prefix")
        assert jg == jr, f"{tag}: JSON differs"
        assert mg == ref.meta(), f"{tag}: meta differs"
        return mg


def run_config(data, k, say, num_bins=None, autotune=False, track_short=True):
    """Both pipelines phase by phase on one workload; returns the GPU context
    (after P4) for further checks."""
    ascii_, off, names = data
    g = cf1g.Cf1g(k=k, num_bins=num_bins or 128)
    g.load_records(ascii=ascii_, rec_off=off)
    target = 0
    if autotune:
        res = g.autotune(reps=1)
        num_bins, target = res["num_bins"], res["p3_target_seg"]
        say(f"autotuner: {num_bins} bins, segment target {target}")
        g.close()
        g = cf1g.Cf1g(k=k, num_bins=num_bins, p3_target_seg=target)
        g.load_records(ascii=ascii_, rec_off=off)
    ref = RefPipeline(k, num_bins=num_bins, threads=THREADS, target_seg=target)
    ref.load_arrays(ascii_, off, names)
    tag = f"k={k}, {len(off) - 1} records, {int(off[-1])} bases, {num_bins} bins"
    say(f"{tag}: oracle on {THREADS} threads")

    h_ref = ref.p1()
    np.testing.assert_array_equal(g.count_minimizer_histogram(), h_ref, err_msg=f"{tag}: P1 histogram")
    b2b_ref, nb_ref = ref.p2()
    b2b, nb = g.partition_minimizers_and_route()
    np.testing.assert_array_equal(b2b, b2b_ref, err_msg=f"{tag}: P2 bucket->bin")
    assert nb == nb_ref
    rw, bstart, recs = g.bins_export()
    gb = sort_records(*gpu_records(rw, bstart, recs))
    rb = sort_records(*ref.bin_records(rw))
    del recs
    assert len(gb[0]) == len(rb[0]), f"{tag}: routed record counts"
    for x, y, what in zip(gb, rb, ("bins", "lengths", "bases")):
        np.testing.assert_array_equal(x, y, err_msg=f"{tag}: routed records ({what})")
    del gb, rb
    say(f"P1 histogram, partition and {int(bstart[-1])} routed records equal")

    d_ref = ref.p3()
    d = g.build_mphf()
    assert d == d_ref, f"{tag}: D"
    lo, hi = g.keys()
    rlo, rhi = ref.keys()
    np.testing.assert_array_equal(lo, rlo, err_msg=f"{tag}: keys (lo)")
    if k > 32:
        np.testing.assert_array_equal(hi, rhi, err_msg=f"{tag}: keys (hi)")
    del lo, hi, rlo, rhi
    np.testing.assert_array_equal(g.mphf_export(), ref.mphf_export(), err_msg=f"{tag}: BBHash structure")
    say(f"D = {d}, keys and MPHF equal")

    ref.p4()
    g.classify_vertices()
    np.testing.assert_array_equal(g.states(), ref.states(), err_msg=f"{tag}: P4 state words")
    say("state words equal")
    meta = compare_p5(g, ref, names, track_short, tag)
    say(f"P5 equal: {meta['unitigs']} unitigs")
    del ref
    return g, d


def test_chunk_edge_cases(progress):
    # records around the 4 Mbp chunk size: one short of it, exactly it, just
    # past it (no k-mer, some k-mers, all k-mers across the edge), two edges
    # with an N run and a lowercase stretch across them; short records between
    C = 4_000_000
    rng = np.random.default_rng(21)
    lens = [C - 1, C, C + 1, C + 29, C + 30, C + 31, C + 100, 2 * C + 5000]
    recs, names = [], []
    for i, n in enumerate(lens):
        s = np.frombuffer(b"ACGT", np.uint8)[rng.integers(0, 4, n)].copy()
        if i == len(lens) - 1:
            s[C - 5:C + 20] = ord("N")
            s[2 * C - 40:2 * C + 40] = np.frombuffer(s[2 * C - 40:2 * C + 40].tobytes().lower(), np.uint8)
        recs.append(s)
        names.append(f"long{i}")
        recs.append(np.frombuffer(b"ACGTTGCAAT" * (i + 3), np.uint8).copy())
        names.append(f"short{i}")
    off = np.zeros(len(recs) + 1, np.uint64)
    off[1:] = np.cumsum([len(r) for r in recs])
    g, d = run_config((np.concatenate(recs), off, names), 31, progress, num_bins=64)
    g.close()


def test_genome_5m_chunk_edge(progress):
    data = synth.genome_5m()
    assert int(data[1][-1]) > 4_000_000
    g, d = run_config(data, 31, progress, num_bins=128)
    g.close()


@pytest.mark.parametrize("k", [31, 21])
def test_transcriptome_100k(transcriptome, k, progress):
    off = transcriptome[1]
    assert len(off) - 1 == 100_000
    assert (np.diff(off.astype(np.int64)) < k).sum() > 1000  # 2% shorter than k
    g, d = run_config(transcriptome, k, progress, num_bins=128, track_short=True)
    g.close()


@pytest.mark.parametrize("k", [31, 63])
def test_genome_300m(genome300m, k, progress):
    off = genome300m[1]
    assert int(off[-1]) == 300_000_000 and len(off) - 1 == 24
    g, d = run_config(genome300m, k, progress, autotune=True)
    g.close()


def canonical_graph(g, names):
    """cf_seg strings and cf_seq tilings with ids replaced by those strings."""
    with tempfile.TemporaryDirectory() as td:
        p = os.path.join(td, "o")
        g.p5_write(p, names=names, threads=THREADS)
        with open(p + ".cf_seg") as f:
            seg = dict(line.split("\t") for line in f.read().splitlines())
        tiles = []
        with open(p + ".cf_seq") as f:
            for line in f.read().splitlines():
                name, toks = line.split("\t")
                tiles.append((name, [t if t.startswith("N") else (seg[t[:-1]], t[-1]) for t in toks.split()]))
    return sorted(seg.values()), sorted(tiles)


def test_learned_index_all_ranks(genome300m, progress):
    ascii_, off, names = genome300m
    g = cf1g.Cf1g(k=31, num_bins=512, p3_target_seg=2048)
    g.load_records(ascii=ascii_, rec_off=off)
    d = g.run_p1_p4()
    lo, _ = g.keys()
    srt = np.sort(lo)
    assert np.all(srt[1:] > srt[:-1]), "distinct keys"
    info = g.lidx_train()
    progress(f"trained on {d} keys: max error {info['max_err_lo']}/{info['max_err_hi']}, "
             f"{info['fallback_leaves']} fallback leaves")
    ranks = g.lidx_lookup(srt)
    np.testing.assert_array_equal(ranks, np.arange(d, dtype=np.uint64), err_msg="learned ranks")
    perm = np.random.default_rng(3).permutation(d)
    np.testing.assert_array_equal(g.lidx_lookup(srt[perm]), perm.astype(np.uint64), err_msg="learned ranks, shuffled order")
    del srt, perm, ranks, lo
    progress(f"all {d} ranks exact")
    want = canonical_graph(g, names)
    g.classify_via(1)
    g.select_index(1)
    got = canonical_graph(g, names)
    assert got == want, "graph through the learned index differs from the BBHash path's"
    progress("graph through the learned index equals the BBHash path's")
    g.close()
