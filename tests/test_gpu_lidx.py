"""GPU learned rank index, checked against the CPU oracle (oracle/lidx_ref.py).

Exact ranks for every key (the sorted order is the map), the key warp's
tables and warped keys, the root records, routing and recorded error bounds
recomputed with the device arithmetic, the
lines byte for byte (both line sizes, overflowing lines, clustered keys whose
runs are searched), forward / backward gradients against double-precision
recomputation, convergence to the closed-form optimum, the gain of training
over endpoint interpolation, the BBHash fallback, and classification / P5
output through the learned index equal (canonicalised) to the BBHash path.
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
    assert info["n_keys"] == D and info["n_leaves"] == (1 << LB) + (1 << RB) and info["n_pieces"] == 1 << RB
    assert info["n_lines"] == -(-D // info["line_keys"]) and info["line_bytes"] == 64 and info["line_keys"] == 108
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
    near = keys[rng.integers(0, D, 20000)] + np.uint64(1)  # next to present keys: same cells
    near = near[~np.isin(near, keys)]
    for q in (absent, near):
        out = g.lidx_lookup(q)
        assert np.all(out < D)  # 0 or some rank, never out of range
    wide = np.array([np.uint64(1) << np.uint64(63), ~np.uint64(0)], np.uint64)  # wider than a 62-bit key
    np.testing.assert_array_equal(g.lidx_lookup(wide), [0, 0])
    rep = np.repeat(keys[:: max(1, D // 50)], 7)
    np.testing.assert_array_equal(g.lidx_lookup(rep), np.searchsorted(keys, rep).astype(np.uint64))


@pytest.mark.parametrize("line_bytes,line_keys", [(128, 0), (64, 0), (128, 64), (64, 400)])
def test_model_and_lines_match_oracle(built, line_bytes, line_keys):
    """Root records, routing, leaf error bounds and every line byte equal the
    oracle's (lines from the sorted keys); 64 keys in 128 bytes leaves slack,
    400 keys in 64 bytes overflow every full line (searched)."""
    g, _, D = built
    info = g.lidx_train(leaf_bits=LB, root_bits=RB, line_bytes=line_bytes, line_keys=line_keys)
    m = g.lidx_export()
    keys, root, leaf, s = m["keys"], m["root"], m["leaf"], m["shift"]
    C = m["line_keys"]
    cum = LR.warp_tables(keys, 2 * K)
    np.testing.assert_array_equal(m["warp"], cum)
    u = LR.warp(keys, cum, 2 * K)
    p, s_o, R = LR.root_pieces(u, RB)
    assert s == s_o and len(root) == R + 1
    lo, leaf0 = LR.root_ranges(p, 1 << LB, D)
    np.testing.assert_array_equal(root["leaf0"], leaf0)
    np.testing.assert_array_equal(root["p"], p)
    assert np.all(root["A"] >= 0)
    j = LR.route(u, root, s)
    assert np.all(np.diff(j) >= 0)
    elo, ehi, _ = LR.error_bounds(u, root, leaf, s)
    np.testing.assert_array_equal(m["err"][:, 0].astype(np.int64), elo)
    np.testing.assert_array_equal(m["err"][:, 1].astype(np.int64), ehi)
    assert info["max_err_lo"] == elo.max() and info["max_err_hi"] == ehi.max()
    want, st = LR.encode_lines(keys, C, m["line_bytes"])
    np.testing.assert_array_equal(m["lines"], want)
    assert info["overflow_lines"] == sum(1 for _, o, _ in st if o)
    assert info["search_keys"] == sum(n for _, _, n in st)
    if line_keys == 400:
        assert info["overflow_lines"] >= len(want) - 1
    rng = np.random.default_rng(1)
    perm = rng.permutation(D)
    np.testing.assert_array_equal(g.lidx_lookup(keys[perm]), perm.astype(np.uint64))


def _clustered_genome(rng):
    """Copies of one motif with substitutions near their ends: k-mers that
    share all but their last bases, in clusters the leaves cannot follow."""
    motif = rng.integers(0, 4, 40)
    seqs = ["".join("ACGT"[c] for c in rng.integers(0, 4, 30000)).encode()]
    for i in range(400):
        cp = motif.copy()
        cp[rng.integers(28, 40)] = rng.integers(0, 4)
        cp[rng.integers(30, 40)] = rng.integers(0, 4)
        seqs.append("".join("ACGT"[c] for c in cp).encode())
    return seqs


def test_walk_across_overflow_lines():
    """Lines of 150 keys do not fit 64 bytes (searched), and one epoch at a
    step size of 10^4 spoils the leaves (errors of many lines): the walk
    crosses runs of overflow lines, searching each one, until it reaches the
    key's own."""
    rng = np.random.default_rng(32)
    g = cf1g.Cf1g(k=K, num_bins=8)
    try:
        g.load_records(_clustered_genome(rng))
        D = g.run_p1_p4()
        info = g.lidx_train(leaf_bits=2, root_bits=1, line_bytes=64, line_keys=150, epochs=1, lr=1.0e4,
                            max_window=1 << 31)
        assert info["fallback_leaves"] == 0
        assert info["overflow_lines"] >= info["n_lines"] // 2
        assert max(info["max_err_lo"], info["max_err_hi"]) > 3 * 150
        keys = g.lidx_export()["keys"]
        perm = rng.permutation(D)
        np.testing.assert_array_equal(g.lidx_lookup(keys[perm]), perm.astype(np.uint64))
    finally:
        g.close()


def test_clustered_keys_search_runs():
    """Clustered k-mers: adjacent keys differ only in low bits, so runs carry
    codes of 7 and resolve by a search of the keys."""
    rng = np.random.default_rng(31)
    g = cf1g.Cf1g(k=K, num_bins=8)
    try:
        g.load_records(_clustered_genome(rng))
        D = g.run_p1_p4()
        info = g.lidx_train(leaf_bits=8, root_bits=4)
        m = g.lidx_export()
        keys = m["keys"]
        want, st = LR.encode_lines(keys, m["line_keys"], m["line_bytes"])
        np.testing.assert_array_equal(m["lines"], want)
        assert info["search_keys"] > 0
        perm = rng.permutation(D)
        np.testing.assert_array_equal(g.lidx_lookup(keys[perm]), perm.astype(np.uint64))
    finally:
        g.close()


@pytest.mark.parametrize("level", [0, 1])
def test_gradients_match_oracle(built, level):
    g, _, D = built
    g.lidx_train(leaf_bits=LB, root_bits=RB)
    m = g.lidx_export()
    keys = LR.warp(m["keys"], m["warp"], 2 * K)
    ns = (1 << RB) if level == 0 else (1 << LB) + (1 << RB)
    rng = np.random.default_rng(7 + level)
    w = rng.standard_normal(ns).astype(np.float32) * np.float32(30)
    c = rng.standard_normal(ns).astype(np.float32) * np.float32(30)
    r = g.lidx_gradients(level, w, c)
    seg = r["seg"]
    if level == 0:
        p, s, _ = LR.root_pieces(keys, RB)
        np.testing.assert_array_equal(seg, p)
        lo, _ = LR.root_ranges(p, 1 << LB, D)
        x, y, sg = LR.level_xy(0, keys, seg, s=s, lo=lo, L=1 << LB)
    else:
        np.testing.assert_array_equal(seg, LR.leaf_starts(keys, m["root"], m["shift"], ns))
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
    keys = LR.warp(m["keys"], m["warp"], 2 * K)
    ns = (1 << LB) + (1 << RB)
    seg = LR.leaf_starts(keys, m["root"], m["shift"], ns).astype(np.uint64)
    x, y, sg = LR.level_xy(1, keys, seg)
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
    assert max(tr["max_err_lo"], tr["max_err_hi"]) <= max(un["max_err_lo"], un["max_err_hi"])


def test_fallback_leaves_stay_exact(built):
    g, _, D = built
    info = g.lidx_train(leaf_bits=LB, root_bits=RB, max_window=3)
    assert info["fallback_leaves"] > 0
    m = g.lidx_export()
    assert np.any(LR.is_fallback(m["leaf"]))
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
