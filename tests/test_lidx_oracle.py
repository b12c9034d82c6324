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


def test_error_bounds_definition():
    rng = np.random.default_rng(4)
    keys = np.sort(rng.choice(np.uint64(1) << np.uint64(30), 500, replace=False).astype(np.uint64))
    L = 8
    q = np.linspace(0, 500, L + 1).astype(np.int64)
    leaf = np.zeros(L, [("base", "<u8"), ("q", "<u4"), ("n", "<u4"), ("S", "<f4"), ("I", "<f4"), ("elo", "<u2"),
                        ("ehi", "<u2"), ("shift", "u1"), ("flags", "u1"), ("pad", "<u2")])
    for j in range(L):
        a, b = q[j], q[j + 1]
        leaf[j]["base"] = keys[a]
        leaf[j]["q"] = a
        leaf[j]["n"] = b - a
        span = float(keys[b - 1] - keys[a]) or 1.0
        leaf[j]["S"] = np.float32((b - a - 1) / span * rng.uniform(0.8, 1.2))
        leaf[j]["I"] = np.float32(rng.uniform(-3, 3))
    elo, ehi, _ = LR.error_bounds(keys, leaf)
    j = np.repeat(np.arange(L), np.diff(q))
    pred = LR.leaf_pred(keys, leaf, j)
    rank = np.arange(500)
    assert np.all(rank >= pred - elo[j]) and np.all(rank <= pred + ehi[j])
    for t in range(L):  # bounds are attained (tight) or zero
        sel = j == t
        e = rank[sel] - pred[sel]
        assert elo[t] == max(0, int((-e).max())) and ehi[t] == max(0, int(e.max()))


def test_route_is_monotone_for_nonnegative_slopes():
    rng = np.random.default_rng(5)
    k = 15
    keys = np.sort(rng.choice(np.uint64(1) << np.uint64(2 * k), 4000, replace=False).astype(np.uint64))
    p, s, R = LR.root_pieces(keys, k, 6)
    L = LR.leaf_count(len(keys), 8)
    lo, hi = LR.root_ranges(p, L, len(keys))
    root = np.zeros(R, [("A", "<f4"), ("B", "<f4"), ("lo", "<u4"), ("hi", "<u4")])
    root["A"] = np.abs(rng.standard_normal(R)).astype(np.float32) * np.float32(2.0 ** -s * 8)
    root["B"] = rng.standard_normal(R).astype(np.float32)
    root["lo"], root["hi"] = lo, hi
    j = LR.route(keys, root, s)
    assert np.all(np.diff(j) >= 0)
    assert j.min() >= 0 and j.max() < L
