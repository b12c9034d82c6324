"""Pins the learned-index oracle (oracle/lidx_ref.py) on the CPU.

fma32 against exact rational arithmetic; the analytic gradient against
central finite differences of the double-precision loss and against the
closed-form optimum; least_squares against numpy.linalg.lstsq; error bounds
against their definition.
"""
import random
from fractions import Fraction

import numpy as np
import pytest

import lidx_ref as LR


def round_f32_exact(v):
    """Nearest float32 to the rational v, ties to even."""
    r = np.float32(float(v))
    best = None
    for cand in (np.nextafter(r, np.float32(-np.inf)), r, np.nextafter(r, np.float32(np.inf))):
        d = abs(Fraction(float(cand)) - v)
        key = (d, int(np.array(cand, np.float32).view(np.uint32)) & 1)
        if best is None or key < best[0]:
            best = (key, cand)
    return np.float32(best[1])


def test_fma32_matches_exact_rounding():
    rng = np.random.default_rng(1)
    a = rng.standard_normal(3000).astype(np.float32) * np.float32(1e3)
    b = rng.standard_normal(3000).astype(np.float32)
    c = rng.standard_normal(3000).astype(np.float32) * np.float32(1e2)
    # adversarial: products whose exact value sits on a float32 midpoint
    # before the tiny addend (double rounding would go wrong)
    t = np.float32(1 + 2.0 ** -12)
    a = np.concatenate([a, np.full(4, t, np.float32)])
    b = np.concatenate([b, np.full(4, t, np.float32)])
    c = np.concatenate([c, np.array([2.0 ** -60, -(2.0 ** -60), 0.0, 2.0 ** -40], np.float32)])
    got = LR.fma32(a, b, c)
    for i in range(len(a)):
        want = round_f32_exact(Fraction(float(a[i])) * Fraction(float(b[i])) + Fraction(float(c[i])))
        assert got[i] == want, (i, a[i], b[i], c[i], got[i], want)


def _segments(rng, n_segs=12):
    sizes = rng.integers(1, 40, n_segs)
    seg = np.concatenate([[0], np.cumsum(sizes)]).astype(np.uint64)
    keys = np.sort(rng.choice(np.uint64(1) << np.uint64(40), int(seg[-1]), replace=False).astype(np.uint64))
    return keys, seg


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(2)
    keys, seg = _segments(rng)
    x, y, sg = LR.level_xy(1, keys, seg)
    ns = len(seg) - 1
    m, isd = LR.stats(x, sg, ns)
    w = rng.standard_normal(ns)
    c = rng.standard_normal(ns) * 5
    n = np.bincount(sg, minlength=ns).astype(np.float64)
    # analytic gradient in double
    z = (x.astype(np.float64) - m[sg]) * isd[sg]
    r = w[sg] * z + c[sg] - y.astype(np.float64)
    gw_a = 2 * np.bincount(sg, weights=r * z, minlength=ns) / n
    gc_a = 2 * np.bincount(sg, weights=r, minlength=ns) / n
    h = 1e-6
    for j in range(ns):
        e = np.zeros(ns)
        e[j] = h
        lw = (LR.loss_double(x, y, sg, ns, m, isd, w + e, c)[j] - LR.loss_double(x, y, sg, ns, m, isd, w - e, c)[j]) / (2 * h) / n[j]
        lc = (LR.loss_double(x, y, sg, ns, m, isd, w, c + e)[j] - LR.loss_double(x, y, sg, ns, m, isd, w, c - e)[j]) / (2 * h) / n[j]
        assert gw_a[j] == pytest.approx(lw, rel=1e-5, abs=1e-6)
        assert gc_a[j] == pytest.approx(lc, rel=1e-5, abs=1e-6)
    # the oracle's FP32-residual gradient agrees with the double one
    gw, gc, _ = LR.gradients(x, y, sg, ns, m.astype(np.float32), isd.astype(np.float32), w, c)
    np.testing.assert_allclose(gw, gw_a, rtol=2e-3, atol=1e-3)
    np.testing.assert_allclose(gc, gc_a, rtol=2e-3, atol=1e-3)


def test_least_squares_and_zero_gradient():
    rng = np.random.default_rng(3)
    keys, seg = _segments(rng)
    x, y, sg = LR.level_xy(1, keys, seg)
    ns = len(seg) - 1
    m, isd = LR.stats(x, sg, ns)
    w, c = LR.least_squares(x, y, sg, ns, m, isd)
    for j in range(ns):
        sel = sg == j
        if sel.sum() < 2:
            continue
        z = (x[sel].astype(np.float64) - m[j]) * isd[j]
        A = np.stack([z, np.ones_like(z)], 1)
        sol = np.linalg.lstsq(A, y[sel].astype(np.float64), rcond=None)[0]
        assert w[j] == pytest.approx(sol[0], rel=1e-7, abs=1e-7)
        assert c[j] == pytest.approx(sol[1], rel=1e-7, abs=1e-7)
    z = (x.astype(np.float64) - m[sg]) * isd[sg]
    r = w[sg] * z + c[sg] - y.astype(np.float64)
    n = np.bincount(sg, minlength=ns)
    g = np.bincount(sg, weights=r * z, minlength=ns) / np.maximum(n, 1)
    assert np.max(np.abs(g)) < 1e-8


def _model(rng, u, rb, lb):
    """Root and leaf records over warped keys u with random non-negative
    slopes (routing and bounds must hold for any such model)."""
    D = len(u)
    p, s, R = LR.root_pieces(u, rb)
    L = LR.leaf_count(D, lb)
    _, leaf0 = LR.root_ranges(p, L, D)
    root = np.zeros(R + 1, LR.ROOT_DTYPE)
    root["A"][:R] = np.abs(rng.standard_normal(R)).astype(np.float32) * np.float32(2.0 ** -s * L / R * 1.5)
    root["B"][:R] = rng.standard_normal(R).astype(np.float32)
    root["leaf0"] = leaf0
    root["p"] = p
    LT = L + R
    leaf = np.zeros(LT, LR.LEAF_DTYPE)
    leaf["S"] = np.float32(D / 2.0 ** 64) * rng.uniform(0.8, 1.2, LT).astype(np.float32)
    leaf["I"] = rng.uniform(-20, 20, LT).astype(np.float32)
    return root, leaf, s, LT


def _warped(rng, n, k=15):
    keys = np.sort(rng.choice(np.uint64(1) << np.uint64(2 * k), n, replace=False).astype(np.uint64))
    cum = LR.warp_tables(keys, 2 * k)
    return keys, LR.warp(keys, cum, 2 * k)


def test_route_is_monotone_and_pieces_own_their_leaves():
    rng = np.random.default_rng(5)
    keys, u = _warped(rng, 4000)
    root, leaf, s, LT = _model(rng, u, 6, 8)
    j = LR.route(u, root, s)
    assert np.all(np.diff(j) >= 0)
    assert j.min() >= 0 and j.max() < LT
    r = (u >> np.uint64(s)).astype(np.int64)
    assert np.all(j >= root["leaf0"][r]) and np.all(j < root["leaf0"][r + 1])
    # a leaf never holds keys of two pieces
    for t in np.unique(j):
        assert len(np.unique(r[j == t])) == 1
    # every piece owns at least one leaf, L + R in all
    assert np.all(np.diff(root["leaf0"].astype(np.int64)) >= 1) and root["leaf0"][-1] == LT


def test_error_bounds_definition():
    rng = np.random.default_rng(4)
    keys, u = _warped(rng, 3000)
    root, leaf, s, LT = _model(rng, u, 5, 6)
    elo, ehi, _ = LR.error_bounds(u, root, leaf, s)
    j = LR.route(u, root, s)
    pred = LR.leaf_pred(u, root, leaf, s, j)
    rank = np.arange(len(u))
    assert np.all(rank >= pred - elo[j]) and np.all(rank <= pred + ehi[j])
    for t in range(LT):  # bounds are attained (tight) or zero
        sel = j == t
        if not sel.any():
            assert elo[t] == 0 and ehi[t] == 0
            continue
        e = rank[sel] - pred[sel]
        assert elo[t] == max(0, int((-e).max())) and ehi[t] == max(0, int(e.max()))
    # the sign bit of S only marks the leaf: the prediction uses |S|
    leaf2 = leaf.copy()
    leaf2["S"] = -leaf2["S"]
    assert np.all(LR.is_fallback(leaf2)) and not np.any(LR.is_fallback(leaf))
    np.testing.assert_array_equal(LR.leaf_pred(u, root, leaf2, s, j), pred)


# ---------------------------------------------------------------- warp
def _warp_exact(key, cum, K2):
    """The interval recurrence in Python integers."""
    lo, w = 0, (1 << 64) - 1
    G = len(cum)
    for g in range(G):
        sym = (key >> (K2 - 8 * (g + 1))) & 255
        a = (w * int(cum[g][sym])) >> 32
        b = w if sym == 255 else (w * int(cum[g][sym + 1])) >> 32
        lo += a
        w = b - a
    rb = K2 - 8 * G
    if rb > 0:
        lo += (w * ((key & ((1 << rb) - 1)) << (64 - rb))) >> 64
    return lo


@pytest.mark.parametrize("k", [31, 21, 15, 5, 3])
def test_warp_integer_recurrence_and_monotone(k):
    rng = np.random.default_rng(20 + k)
    K2 = 2 * k
    n = min(3000, 1 << K2)
    keys = np.sort(rng.choice(np.uint64(1) << np.uint64(K2), n, replace=False).astype(np.uint64))
    # skew the byte counts (GC-poor genome-like keys) through the tables
    cum = LR.warp_tables(keys[: n // 3], K2)
    edge = np.array([0, (1 << K2) - 1, (1 << (K2 - 1)), (1 << (K2 - 1)) - 1], np.uint64)
    allk = np.unique(np.concatenate([keys, edge]))
    u = LR.warp(allk, cum, K2)
    for i in range(0, len(allk), 37):
        assert int(u[i]) == _warp_exact(int(allk[i]), cum, K2)
    for e in edge:
        assert int(LR.warp(np.array([e], np.uint64), cum, K2)[0]) == _warp_exact(int(e), cum, K2)
    assert np.all(np.diff(u.astype(object)) >= 0)  # non-decreasing in the key


def test_warp_of_uniform_counts_is_the_scaled_key():
    """Equal byte counts make every interval split evenly: u is the key
    scaled to 64 bits, up to the floors (a few units per group)."""
    K2 = 62
    i = np.arange(1 << 16, dtype=np.uint64)
    # every group's byte takes each value 256 times (odd multipliers permute bytes)
    g0, g1, g2 = i & np.uint64(255), i >> np.uint64(8), (i * np.uint64(37)) & np.uint64(255)
    rest = (i * np.uint64(2654435761)) & np.uint64((1 << 38) - 1)
    keys = np.unique((g0 << np.uint64(54)) | (g1 << np.uint64(46)) | (g2 << np.uint64(38)) | rest)
    assert len(keys) == 1 << 16
    cum = LR.warp_tables(keys, K2)
    assert cum.shape == (LR.WARP_MAX_GROUPS, 256) == (3, 256)
    np.testing.assert_array_equal(cum, np.tile(np.arange(256, dtype=np.uint32) << np.uint32(24), (3, 1)))
    u = LR.warp(keys, cum, K2).astype(object)
    want = [int(x) << 2 for x in keys]
    assert max(abs(a - b) for a, b in zip(u, want)) < (1 << 34)


def test_piece_x_keeps_the_top_32_bits():
    """piece_x equals the exact FP32 conversion of the top 32 bits of x, and
    differs from the exact conversion of x by less than 2^(s-32) + an ulp."""
    rng = np.random.default_rng(22)
    s = 54
    u = rng.integers(0, 1 << 63, 50000, dtype=np.uint64) * np.uint64(2) + rng.integers(0, 2, 50000, dtype=np.uint64)
    x = LR.piece_x(u, s).astype(np.float64)
    low = (u & np.uint64((1 << s) - 1)).astype(object)
    for a, b in zip(x[:2000], low[:2000]):
        top = b >> (s - 32)
        assert a == float(np.float32(top)) * 2.0 ** (s - 32)
        assert abs(a - b) <= 2.0 ** (s - 32) + abs(a) * 2.0 ** -24
    assert np.all(LR.piece_x(np.array([0, 1 << s, (1 << s) + 5], np.uint64), s) == 0)


def test_warp_flattens_skewed_keys():
    """Keys whose bases follow a 0.35/0.15/0.15/0.35 composition: raw keys
    cluster, warped keys spread evenly (Kolmogorov distance to uniform)."""
    rng = np.random.default_rng(21)
    n = 200000
    bases = rng.choice(4, size=(n, 31), p=[0.35, 0.15, 0.15, 0.35])
    keys = np.zeros(n, np.uint64)
    for b in range(31):
        keys = (keys << np.uint64(2)) | bases[:, b].astype(np.uint64)
    keys = np.unique(keys)
    cum = LR.warp_tables(keys, 62)
    u = np.sort(LR.warp(keys, cum, 62))
    ecdf = np.arange(len(keys)) / len(keys)
    d_raw = np.max(np.abs(keys.astype(np.float64) / 2.0 ** 62 - ecdf))
    d_warp = np.max(np.abs(u.astype(np.float64) / 2.0 ** 64 - ecdf))
    assert d_raw > 0.05 and d_warp < 0.01


# ---------------------------------------------------------------- lines
def test_line_format_by_hand():
    """Keys 16, 17, 20, 40 in one line (C = 4): rel 0, 1, 4, 24; t = 4 gives
    cells [0, 0, 0, 1] (12 bits: 4 + 2 cells + 2 codes of 3), so the unary is
    1110 10, and the codes are t - bitlen(1 ^ 0) = 3 and t - bitlen(4 ^ 1) = 1."""
    keys = np.array([16, 17, 20, 40], np.uint64)
    assert LR.line_choice([16, 17, 20, 40], 64) == (4, 12, 2, 3)
    lines, st = LR.encode_lines(keys, 4, 64)
    w = [int(x) for x in lines[0].view(np.uint64)]
    assert w[0] == 16  # k0, no overflow
    assert w[1] & 63 == 4 and (w[1] >> 6) & 1023 == 6  # t, ulen
    assert (w[1] >> 16) & 255 == 0  # no sentinel: prefix 0
    unary = [(w[1] >> (24 + b)) & 1 for b in range(6)]
    assert unary == [1, 1, 1, 0, 1, 0]
    codes = (w[1] >> 30) & 0x3F
    assert codes == 3 | (1 << 3)
    assert st[0] == (12, False, 0)
    for r, q in enumerate(keys):
        assert LR.line_lookup(keys, lines, 4, int(q), 0) == r
    # absent keys in an occupied cell resolve to some member, never out of range
    for q in (18, 19, 21, 33, 41, 15):
        got = LR.line_lookup(keys, lines, 4, q, 0)
        assert got is None or 0 <= got < 4


def test_line_takes_4bit_codes_for_a_close_pair():
    """rel 0, 1, 2000, 4000: t = 10 is best for 3-bit codes (11 bits), but
    the pair (0, 1) differs only in bit 0 (code 10 - 1 = 9 > 7, a search);
    4-bit codes fit (12 bits), so the line takes them and the pair resolves
    by its code."""
    b = 1 << 20
    ks = [b, b + 1, b + 2000, b + 4000]
    assert LR.line_choice(ks, 64) == (10, 12, 1, 4)
    keys = np.array(ks, np.uint64)
    lines, st = LR.encode_lines(keys, 4, 64)
    w = [int(x) for x in lines[0].view(np.uint64)]
    assert (w[0] >> 63) & 1 == 1 and (w[0] >> 62) & 1 == 0
    assert w[1] & 63 == 10 and (w[1] >> 6) & 1023 == 8
    assert (w[1] >> (24 + 8)) & 15 == 9  # the one code, 4 bits
    assert st[0] == (12, False, 0)
    for r, q in enumerate(ks):
        assert LR.decode_line(lines[0], q, 0, 4, 1) == ("found", r)


def _check_all(keys, C, line_bytes, starts=(0,)):
    lines, st = LR.encode_lines(keys, C, line_bytes)
    D = len(keys)
    for r in range(D):
        true = r // C
        for d in starts:
            got = LR.line_lookup(keys, lines, C, int(keys[r]), true + d)
            assert got == r, (r, d, got)
    return lines, st


@pytest.mark.parametrize("line_bytes,C", [(64, 40), (128, 96), (128, 256)])
def test_lines_uniform_keys_from_any_start(line_bytes, C):
    rng = np.random.default_rng(11)
    keys = np.sort(rng.choice(np.uint64(1) << np.uint64(62), 3 * C + 17, replace=False).astype(np.uint64))
    lines, st = _check_all(keys, C, line_bytes, starts=(0, -2, -1, 1, 2))
    assert not any(o for _, o, _ in st)


def test_lines_clustered_keys_need_search():
    """Clusters of keys that share all but their low bits: long runs and
    codes of 7 (searched); sizes and tiny D."""
    rng = np.random.default_rng(12)
    base = rng.choice(np.uint64(1) << np.uint64(62), 60, replace=False).astype(np.uint64)
    clus = []
    for b in base:
        bits = int(rng.integers(5, 12))
        clus.append(b + rng.choice(np.uint64(1) << np.uint64(bits), int(rng.integers(2, 30)), replace=False).astype(np.uint64))
    keys = np.unique(np.concatenate([base] + clus))
    lines, st = _check_all(keys, 64, 128, starts=(0, 1, -1))
    assert sum(n for _, _, n in st) > 0  # some runs resolve by search
    for D in (1, 2, 63, 64, 65, 129):
        _check_all(keys[:D], 64, 128, starts=(0, 1))


def test_lines_overflow_search():
    rng = np.random.default_rng(13)
    keys = np.sort(rng.choice(np.uint64(1) << np.uint64(62), 700, replace=False).astype(np.uint64))
    lines, st = _check_all(keys, 200, 64, starts=(0, -3, -2, -1, 1, 2, 3))  # 200 keys never fit 432 bits
    assert all(o for _, o, _ in st[:-1])  # the last line holds 100 keys


def test_sentinel_prefix_by_hand():
    """Line 0 of keys 64, 70, 77, 79 (C = 3) with sentinel 79: rel 0, 6, 13,
    15 (pair bit lengths 3, 4, 2), tb = 1 and t = 2 needs 4 + 3 + 1 + 3 = 11
    bits against 12, 12, 20 and 14 for t = 1, 3, 0 and 4, so cells are
    [0, 1, 3, 3]; the sentinel's offset in its cell is 3, prefix 3 << 6 = 192
    (t < 8: exact). q = 77 (offset 1, prefix 64 < 192) resolves in the run
    without the sentinel (one key, rank 2); q = 79 ties (searched, found at
    rank 3). A query above the sentinel in its cell goes right."""
    keys = np.array([64, 70, 77, 79, 200], np.uint64)
    ks = [64, 70, 77, 79]
    t, bits, coll, cw = LR.line_choice(ks, 64)
    assert (t, bits, cw) == (2, 11, 3)
    lines, st = LR.encode_lines(keys, 3, 64)
    w = [int(x) for x in lines[0].view(np.uint64)]
    assert (w[1] >> 16) & 255 == LR.sentinel_prefix(15, 2) == 192
    assert LR.decode_line(lines[0], 77, 0, 3, 2) == ("found", 2)
    assert LR.decode_line(lines[0], 79, 0, 3, 2) == ("search", 2, 3, True)
    assert LR.line_lookup(keys, lines, 3, 79, 0) == 3
    # rel 0, 100, 519, 812 (sentinel), then 912 in line 1: t = 8 (11 bits
    # against 14 and 12 for t = 7 and 9), cells 0, 0, 2, 3; 912 lies in the
    # sentinel's cell 3 with offset 144 > 44 (exact prefixes at t = 8): right
    assert LR.sentinel_prefix(300, 9) == 150 and LR.sentinel_prefix(812, 8) == 44
    base = 1 << 30
    keys = np.array([base, base + 100, base + 519, base + 812, base + 912], np.uint64)
    t, bits, coll, cw = LR.line_choice([int(x) for x in keys[:4]], 64)
    assert (t, bits) == (8, 11)
    lines, st = LR.encode_lines(keys, 3, 64)
    w = [int(x) for x in lines[0].view(np.uint64)]
    assert (w[1] >> 16) & 255 == 44
    assert LR.decode_line(lines[0], int(keys[4]), 0, 3, 2) == ("right",)
    for r, q in enumerate(keys):
        assert LR.line_lookup(keys, lines, 3, int(q), 0) == r


def test_lines_sentinel_cell_shared():
    """The next line's first key shares a cell with this line's last keys:
    keys of both lines resolve (the run with the sentinel: prefixes decide,
    ties are searched)."""
    keys = np.array([1000, 1001, 1002, 1003, 1004, 1005, 1006, 5000, 5001], np.uint64)
    _check_all(keys, 4, 64, starts=(0, 1, -1))
    keys = np.array([1 << 40, (1 << 40) + 3, (1 << 40) + 5, (1 << 40) + 6, (1 << 40) + 7, (1 << 41)], np.uint64)
    _check_all(keys, 3, 64, starts=(0, 1, -1))


def test_magic_division():
    """line = pred / C by multiply-high, as the device computes it."""
    rng = np.random.default_rng(14)
    for C in (2, 3, 7, 100, 104, 112, 255, 256, 257, 4096):
        l = (C - 1).bit_length()
        M = ((((1 << l) - C) << 32) // C + 1) & 0xFFFFFFFF
        ns = np.concatenate([rng.integers(0, 1 << 32, 20000, dtype=np.uint64), np.arange(0, 3 * C, dtype=np.uint64),
                             np.array([(1 << 32) - 1, (1 << 32) - 2], np.uint64)])
        for n in ns:
            n = int(n)
            t = (n * M) >> 32
            assert (t + ((n - t) >> 1)) >> (l - 1) == n // C
