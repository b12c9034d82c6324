"""cf1g_p1_stream: P1 over host records while they stream to the device
(host-packed chunks through pinned staging, the scan following the copies)
gives the products of cf1g_load_records + P1..P4 exactly, for inputs of one
chunk and of many, chunk edges of long records, odd records and any number
of packing threads."""
import numpy as np
import pytest

import cf1g
from cf1g import synth
from gpu_cases import edge_records, random_genome

pytestmark = pytest.mark.gpu


def products(g):
    d = g.build_mphf()
    g.classify_vertices()
    lo, hi = g.keys()
    return d, lo, hi, g.mphf_export(), g.states(), g.annotate()


def compare(k, bins, ascii_, off, threads=(0,)):
    g = cf1g.Cf1g(k=k, num_bins=bins)
    g.load_records(ascii=ascii_, rec_off=off)
    h = g.count_minimizer_histogram()
    b2b, nb = g.partition_minimizers_and_route()
    want = (h, b2b, nb) + products(g)
    g.close()
    for th in threads:
        s = cf1g.Cf1g(k=k, num_bins=bins)
        hs = s.p1_stream(ascii=ascii_, rec_off=off, threads=th)
        b2bs, nbs = s.partition_minimizers_and_route()
        got = (hs, b2bs, nbs) + products(s)
        assert got[2] == want[2] and got[3] == want[3]
        for x, y in zip(got[:2] + got[4:], want[:2] + want[4:]):
            np.testing.assert_array_equal(x, y, err_msg=f"k={k} threads={th}")
        np.testing.assert_array_equal(s.stream_export()[0], g_stream(k, bins, ascii_, off)[0])
        s.close()


def g_stream(k, bins, ascii_, off):
    g = cf1g.Cf1g(k=k, num_bins=bins)
    g.load_records(ascii=ascii_, rec_off=off)
    g.count_minimizer_histogram()
    out = g.stream_export()
    g.close()
    return out


def arrays(seqs):
    off = np.zeros(len(seqs) + 1, np.uint64)
    off[1:] = np.cumsum([len(s) for s in seqs])
    a = np.frombuffer(b"".join(seqs), np.uint8) if seqs and off[-1] else np.zeros(1, np.uint8)
    return a, off


@pytest.mark.parametrize("k", [5, 31, 63])
def test_stream_small_inputs(k):
    compare(k, 8, *arrays(random_genome(seed=400 + k, n_rec=5, max_len=4000) + edge_records()), threads=(1, 3, 0))


def test_stream_empty_and_short():
    compare(31, 4, *arrays([]))
    compare(31, 4, *arrays([b"ACGT", b"", b"GGGG"]))


def test_stream_many_chunks_and_edges():
    # 45 Mbp: several 16.8 M-position chunks, records across chunk boundaries
    # and longer than the 4 Mbp scan chunks of P1 / P2
    a, off, _ = synth.genome_300m(scale=0.15)
    compare(31, 64, a, off, threads=(16, 5))


def test_stream_offset_records():
    # rec_off need not start at 0: the first record starts at byte 7
    seqs = random_genome(seed=5, n_rec=3, max_len=3000)
    a, off = arrays(seqs)
    pad = np.concatenate([np.frombuffer(b"NNNNNNN", np.uint8), a])
    g = cf1g.Cf1g(k=31, num_bins=8)
    g.load_records(ascii=a, rec_off=off)
    h = g.count_minimizer_histogram()
    g.close()
    s = cf1g.Cf1g(k=31, num_bins=8)
    np.testing.assert_array_equal(s.p1_stream(ascii=pad, rec_off=off + np.uint64(7)), h)
    s.close()
