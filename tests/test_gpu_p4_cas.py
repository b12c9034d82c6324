"""GPU P4 in the reference's form (cf1g_p4_classify_cas): one MPHF lookup and
one 64-bit CAS join per k-mer occurrence, in a persistent work-stealing
launch. Its state words must equal the oracle's (which runs the reference's
CAS classifier, src/classify.rs:25-601, over the same MPHF) and the
mask-join P4's, for every chunk size and grid shape, with and without forced
CAS retries (compare_exchange_weak may fail spuriously,
src/state_vector.rs:47-70)."""
import numpy as np
import pytest

import cf1g
from gpu_cases import edge_records, random_genome
from oracle import RefPipeline

pytestmark = pytest.mark.gpu


def oracle_states(seqs, k, num_bins):
    ref = RefPipeline(k, num_bins=num_bins, threads=4)
    ref.load(seqs)
    ref.run_p1_p4()
    return ref.states()


def cas_case(seqs, k, num_bins=16, chunks=((0, 0),), force_retry=(False,)):
    want = oracle_states(seqs, k, num_bins)
    g = cf1g.Cf1g(k=k, num_bins=num_bins)
    g.load_records(seqs)
    g.run_p1_p4()
    joined = g.states()
    np.testing.assert_array_equal(joined, want, err_msg="mask-join P4 vs oracle")
    W = g.info()["n_kmers"]
    g.close()
    for chunk, wpc in chunks:
        g = cf1g.Cf1g(k=k, num_bins=num_bins, p4_chunk=chunk, p4_wgs_per_cu=wpc)
        g.load_records(seqs)
        g.count_minimizer_histogram(want=False)
        g.partition_minimizers_and_route()
        g.build_mphf()
        for fr in force_retry:
            st = g.classify_vertices_cas(force_retry=fr)
            assert st["occurrences"] == W
            np.testing.assert_array_equal(g.states(), want, err_msg=f"CAS P4 vs oracle (chunk {chunk}, wgs {wpc}, retry {fr})")
            assert st["cas_issued"] <= W + st["cas_failed"]
            if fr and st["cas_issued"]:
                # every occurrence that changes a state fails its first CAS once
                assert 2 * st["cas_failed"] >= st["cas_issued"]
        g.close()


@pytest.mark.parametrize("k", [3, 9, 21, 31, 33, 63])
def test_cas_random_genome(k):
    cas_case(random_genome(seed=300 + k, n_rec=5, max_len=4000), k, force_retry=(False, True))


@pytest.mark.parametrize("k", [5, 31, 63])
def test_cas_edge_records(k):
    cas_case(edge_records(), k, num_bins=8, force_retry=(False, True))


def test_cas_chunk_and_grid_shapes():
    seqs = random_genome(seed=41, n_rec=6, max_len=20000)
    cas_case(seqs, 31, chunks=((256, 1), (1000, 3), (8192, 8), (65536, 32), (1 << 24, 2)))


def test_cas_contention():
    # many occurrences of few k-mers: lanes race on the same state words
    seqs = [b"A" * 60000, b"ACGT" * 20000, b"ACGTTGCA" * 5000, b"C" * 3000 + b"N" + b"G" * 3000]
    cas_case(seqs, 31, num_bins=4, chunks=((256, 32), (4096, 8)), force_retry=(False, True))


def test_cas_empty_and_order():
    g = cf1g.Cf1g(k=31, num_bins=4)
    with pytest.raises(cf1g.Cf1gError) as e:
        g.classify_vertices_cas()
    assert e.value.code == -2
    g.load_records([b"", b"ACG"])
    g.run_p1_p4()
    st = g.classify_vertices_cas()
    assert st["occurrences"] == 0 and st["cas_issued"] == 0
    g.close()
    g = cf1g.Cf1g(k=31, num_bins=4, p4_chunk=100)
    g.load_records([b"ACGT" * 100])
    g.run_p1_p4()
    with pytest.raises(cf1g.Cf1gError) as e:
        g.classify_vertices_cas()
    assert e.value.code == -1
    g.close()
