"""GPU learned rank index, checked against the CPU oracle (oracle/lidx_ref.py).

Exact ranks for every key (the sorted order is the map), the routing and the
recorded error bounds recomputed with the device arithmetic, forward /
backward gradients against double-precision recomputation, convergence to
the closed-form optimum, the gain of training over endpoint interpolation,
the BBHash fallback, and classification / P5 output through the learned
index equal (canonicalised) to the BBHash path.
"""
import os
import tempfile

import numpy as np
import pytest

import cf1g
import lidx_ref as LR
from gpu_cases import random_genome
from oracle import RefPipeline

pytestmark = pytest.mark.gpu

K = 31
LB, RB = 10, 6  # 1024 leaves, 64 root pieces for the test inputs


@pytest.fixture(scope="module")
def built():
    seqs = random_genome(seed=21, n_rec=4, max_len=40000)
    g = cf1g.Cf1g(k=K, num_bins=16)
    g.load_records(seqs)
    D = g.run_p1_p4()
    yield g, seqs, D
    g.close()


def sorted_keys(g):
    lo, _ = g.keys()
    return np.sort(lo)


def test_sorted_keys_and_exact_ranks(built):
    g, _, D = built
    info = g.lidx_train(leaf_bits=LB, root_bits=RB)
    m = g.lidx_export()
    keys = sorted_keys(g)
    np.testing.assert_array_equal(m["keys"], keys)
    assert info["n_keys"] == D and info["n_leaves"] == 1 << LB and info["n_pieces"] == 1 << RB
    rng = np.random.default_rng(0)
    perm = rng.permutation(D)
    out = g.lidx_lookup(keys[perm])
    np.testing.assert_array_equal(out, perm.astype(np.uint64))


def test_absent_and_repeated_queries(built):
    g, _, D = built
    g.lidx_train(leaf_bits=LB, root_bits=RB)
    keys = sorted_keys(g)
    rng = np.random.default_rng(3)
    absent = rng.integers(0, 1 << 62, 20000, dtype=np.uint64)
    absent = absent[~np.isin(absent, keys)]
    out = g.lidx_lookup(absent)
    assert np.all(out < D)  # 0 or some rank, never out of range
    wide = np.array([np.uint64(1) << np.uint64(63), ~np.uint64(0)], np.uint64)  # wider than a 62-bit key
    np.testing.assert_array_equal(g.lidx_lookup(wide), [0, 0])
    rep = np.repeat(keys[:: max(1, D // 50)], 7)
    np.testing.assert_array_equal(g.lidx_lookup(rep), np.searchsorted(keys, rep).astype(np.uint64))


def test_model_matches_oracle(built):
    g, _, D = built
    g.lidx_train(leaf_bits=LB, root_bits=RB)
    m = g.lidx_export()
    keys, root, leaf, s = m["keys"], m["root"], m["leaf"], m["shift"]
    # root ranges
    p, s_o, R = LR.root_pieces(keys, K, RB)
    assert s == s_o and len(root) == R
    lo, hi = LR.root_ranges(p, 1 << LB, D)
    np.testing.assert_array_equal(root["lo"], lo)
    np.testing.assert_array_equal(root["hi"], hi)
    assert np.all(root["A"] >= 0)
    # routing: leaf boundaries are where the oracle routes the sorted keys
    j = LR.route(keys, root, s)
    assert np.all(np.diff(j) >= 0)
    q = np.searchsorted(j, np.arange(1 << LB), side="left")
    np.testing.assert_array_equal(leaf["q"].astype(np.int64), q)
    # recorded error bounds are the exact extremes of the lookup arithmetic
    elo, ehi, _ = LR.error_bounds(keys, leaf)
    np.testing.assert_array_equal(leaf["elo"].astype(np.int64), np.minimum(elo, 65535))
    np.testing.assert_array_equal(leaf["ehi"].astype(np.int64), np.minimum(ehi, 65535))


@pytest.mark.parametrize("level", [0, 1])
def test_gradients_match_oracle(built, level):
    g, _, D = built
    g.lidx_train(leaf_bits=LB, root_bits=RB)
    m = g.lidx_export()
    keys = m["keys"]
    ns = (1 << RB) if level == 0 else (1 << LB)
    rng = np.random.default_rng(7 + level)
    w = rng.standard_normal(ns).astype(np.float32) * np.float32(30)
    c = rng.standard_normal(ns).astype(np.float32) * np.float32(30)
    r = g.lidx_gradients(level, w, c)
    seg = r["seg"]
    if level == 0:
        p, s, _ = LR.root_pieces(keys, K, RB)
        np.testing.assert_array_equal(seg, p)
        lo, _ = LR.root_ranges(p, 1 << LB, D)
        x, y, sg = LR.level_xy(0, keys, seg, s=s, lo=lo, L=1 << LB)
    else:
        x, y, sg = LR.level_xy(1, keys, seg)
    mu, isd = LR.stats(x, sg, ns)
    np.testing.assert_allclose(r["mu"], mu.astype(np.float32), rtol=1e-6, atol=1e-3)
    np.testing.assert_allclose(r["isd"], isd.astype(np.float32), rtol=1e-5, atol=0)
    gw, gc, sq = LR.gradients(x, y, sg, ns, r["mu"], r["isd"], w, c)
    np.testing.assert_allclose(r["gw"], gw, rtol=1e-9, atol=1e-9)
    np.testing.assert_allclose(r["gc"], gc, rtol=1e-9, atol=1e-9)
    np.testing.assert_allclose(r["sq"], sq, rtol=1e-9, atol=1e-6)


def test_training_reaches_least_squares(built):
    g, _, D = built
    info = g.lidx_train(leaf_bits=LB, root_bits=RB, epochs=4)
    m = g.lidx_export()
    keys = m["keys"]
    seg = np.concatenate([m["leaf"]["q"].astype(np.uint64), [np.uint64(D)]])
    x, y, sg = LR.level_xy(1, keys, seg)
    ns = 1 << LB
    mu, isd = LR.stats(x, sg, ns)
    w, c = LR.least_squares(x, y, sg, ns, mu.astype(np.float32), isd.astype(np.float32))
    best = LR.loss_double(x, y, sg, ns, mu.astype(np.float32), isd.astype(np.float32), w, c).sum() / D
    assert info["leaf_loss"][1] <= info["leaf_loss"][0]
    assert info["leaf_loss"][1] == pytest.approx(best, rel=1e-3, abs=1e-3)


def test_training_beats_endpoint_interpolation(built):
    g, _, D = built
    un = g.lidx_train(leaf_bits=LB, root_bits=RB, untrained=True)
    keys = sorted_keys(g)
    np.testing.assert_array_equal(g.lidx_lookup(keys), np.arange(D, dtype=np.uint64))
    tr = g.lidx_train(leaf_bits=LB, root_bits=RB)
    np.testing.assert_array_equal(g.lidx_lookup(keys), np.arange(D, dtype=np.uint64))
    assert tr["leaf_loss"][1] < 0.8 * un["leaf_loss"][1]
    assert tr["mean_abs_err"] < un["mean_abs_err"]
    assert tr["mean_window"] < un["mean_window"]


def test_fallback_leaves_stay_exact(built):
    g, _, D = built
    info = g.lidx_train(leaf_bits=LB, root_bits=RB, max_window=3)
    assert info["fallback_leaves"] > 0
    keys = sorted_keys(g)
    np.testing.assert_array_equal(g.lidx_lookup(keys), np.arange(D, dtype=np.uint64))


def test_classify_through_bbhash_equals_p4(built):
    g, _, _ = built
    s4 = g.states().copy()
    g.classify_via(0)
    np.testing.assert_array_equal(g.states(), s4)


def _canon(prefix):
    seg = {}
    for line in open(prefix + ".cf_seg").read().splitlines():
        i, s = line.split("\t")
        seg[i] = s
    tiles = []
    for line in open(prefix + ".cf_seq").read().splitlines():
        name, rest = line.split("\t", 1)
        toks = []
        for t in rest.split(" "):
            if t and t[-1] in "+-" and t[:-1] in seg:
                toks.append(seg[t[:-1]] + t[-1])
            else:
                toks.append(t)
        tiles.append(name + "\t" + " ".join(toks))
    return sorted(seg.values()), sorted(tiles)


def test_graph_output_through_learned_index(built):
    g, seqs, D = built
    names = [f"r{i}" for i in range(len(seqs))]
    g.lidx_train()  # default shape (capped by D)
    g.classify_via(1)
    # states in rank order equal the BBHash states of the same k-mers
    keys = sorted_keys(g)
    g.select_index(0)
    sb = g.states()
    v = g.mphf_lookup(keys)
    g.select_index(1)
    sl = g.states()
    np.testing.assert_array_equal(cf1g.state_of(sl, np.arange(D)), cf1g.state_of(sb, v))
    with tempfile.TemporaryDirectory() as td:
        pl, pb = os.path.join(td, "l"), os.path.join(td, "b")
        ml = g.p5_write(pl, names=names, threads=4)
        g.select_index(0)
        mb = g.p5_write(pb, names=names, threads=4)
        assert ml == mb
        assert _canon(pl) == _canon(pb)
        # and the BBHash path equals the oracle's output
        ref = RefPipeline(K, num_bins=16, threads=4)
        ref.load(seqs, names)
        ref.run_p1_p4()
        pr = os.path.join(td, "r")
        ref.p5(pr)
        for ext in (".cf_seg", ".cf_seq"):
            assert sorted(open(pb + ext).read().splitlines()) == sorted(open(pr + ext).read().splitlines())
