"""Phase launches replayed from hipGraphs: every replay, with eager work and
synchronisations between replays, equals the eager (graph-free) run and the
oracle; new records of another size are recaptured, new records of the same
size reuse the graphs."""
import random

import numpy as np
import pytest

import cf1g
from conftest import rand_seq
from gpu_cases import random_genome
from oracle import RefPipeline

pytestmark = pytest.mark.gpu


def oracle(seqs, k, num_bins):
    ref = RefPipeline(k, num_bins=num_bins, threads=4)
    ref.load(seqs)
    d = ref.run_p1_p4()
    return d, ref.mphf_export(), ref.states()


def products(g):
    return g.info()["n_distinct"], g.mphf_export(), g.states()


def same(a, b, what):
    assert a[0] == b[0], what
    np.testing.assert_array_equal(a[1], b[1], err_msg=what)
    np.testing.assert_array_equal(a[2], b[2], err_msg=what)


@pytest.mark.parametrize("k", [21, 31, 63])
def test_graph_replays_match_eager_and_oracle(k):
    seqs = random_genome(seed=500 + k, n_rec=5, max_len=15000)
    want = oracle(seqs, k, 16)
    eager = cf1g.Cf1g(k=k, num_bins=16)
    eager.set_graphs(False)
    eager.load_records(seqs)
    eager.run_p1_p4()
    same(products(eager), want, "eager")
    assert eager.info()["graph_captures"] == 0
    eager.close()

    g = cf1g.Cf1g(k=k, num_bins=16)
    g.load_records(seqs)
    for rep in range(4):
        g.run_p1_p4()
        same(products(g), want, f"replay {rep}")
        # eager work between replays: lookups, the CAS classifier, a sync
        lo, hi = g.keys()
        assert g.mphf_lookup(lo, hi).tolist() == list(range(want[0]))
        g.classify_vertices_cas()
        same(products(g), want, f"CAS after replay {rep}")
        g.sync()
    # captured on the first run; captured again once after the eager work's
    # first allocations (any new device buffer invalidates the graphs), then
    # replayed
    info = g.info()
    assert info["graph_captures"] > 0
    assert info["graph_replays"] >= info["graph_captures"]
    caps = info["graph_captures"]
    g.run_p1_p4()
    same(products(g), want, "last replay")
    assert g.info()["graph_captures"] == caps
    g.close()


def test_graphs_follow_new_records():
    r = random.Random(3)
    a = [rand_seq(r, 9000).encode(), rand_seq(r, 4000).encode()]
    b = [rand_seq(r, 9000).encode(), rand_seq(r, 4000).encode()]  # same sizes, other bases
    c = [rand_seq(r, 30000).encode()]                              # other sizes
    g = cf1g.Cf1g(k=31, num_bins=8)
    for seqs in (a, b, c, a):
        g.load_records(seqs)
        g.run_p1_p4()
        same(products(g), oracle(seqs, 31, 8), "records")
    g.close()


def test_profiling_runs_eagerly():
    seqs = random_genome(seed=8, n_rec=3, max_len=8000)
    g = cf1g.Cf1g(k=31, num_bins=8)
    g.load_records(seqs)
    g.run_p1_p4()
    caps = g.info()["graph_captures"]
    g.set_profiling(True)
    g.kernel_times(reset=True)
    g.run_p1_p4()
    kt = g.kernel_times(reset=True)
    g.set_profiling(False)
    assert "p3_dedup" in kt and "p1_scan" in kt
    assert g.info()["graph_captures"] == caps
    same(products(g), oracle(seqs, 31, 8), "profiled")
    g.close()
