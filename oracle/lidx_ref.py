"""CPU oracle for the learned rank index (numpy).

TEST INFRASTRUCTURE ONLY: imported by tests/ (and nothing on the product
path). The learned index has no counterpart in the reference; it is an
alternative to the MPHF lookup of src/mphf.rs:268-276 whose values are the
ranks of the keys in the globally sorted distinct k-mers.

What this file restates, for checking the GPU kernels:
  * fma32: the FP32 fused multiply-add of the device, emulated exactly
    (float64 product, TwoSum error term, correct tie handling);
  * the training objective: per segment, inputs x, targets y,
    standardisation z = (x - mu) * isd (FP32), prediction fma(w, z, c),
    residual f - y (FP32) and the MSE gradient (2/n) sum r (z, 1) accumulated
    in double — the forward / backward kernels' arithmetic, key by key;
  * least_squares: the closed-form optimum the training descends towards;
  * route / leaf_pred / error_bounds: the lookup arithmetic (root routing,
    leaf prediction) and the recorded per-leaf error bounds.
Pinning: fma32 against exact rational arithmetic, the gradient against
finite differences of the loss and against the closed-form optimum (zero
gradient there), least_squares against numpy.linalg.lstsq (tests/test_lidx_oracle.py).
"""
import numpy as np


def fma32(a, b, c):
    """Correctly rounded float32 a*b + c (vectorised)."""
    a = np.asarray(a, np.float32)
    b = np.asarray(b, np.float32)
    c = np.asarray(c, np.float32)
    p = a.astype(np.float64) * b.astype(np.float64)  # exact: 24 + 24 bits
    c64 = c.astype(np.float64)
    s = p + c64
    bb = s - p
    e = (p - (s - bb)) + (c64 - bb)  # s + e == p + c exactly
    r = s.astype(np.float32)
    r64 = r.astype(np.float64)
    toward = np.where(s > r64, np.float32(np.inf), np.float32(-np.inf)).astype(np.float32)
    other = np.nextafter(r, toward)
    mid = (r64 + other.astype(np.float64)) * 0.5  # exact: adjacent floats
    is_mid = (s == mid) & (e != 0)
    up = np.maximum(r, other)
    dn = np.minimum(r, other)
    return np.where(is_mid, np.where(e > 0, up, dn), r).astype(np.float32)


# ---------------------------------------------------------------- segments
def root_pieces(keys, k, root_bits):
    """Piece starts p[0..R] (first rank with key >= r << s) and the leaf
    ranges lo_r / hi_r; also returns s, R, L."""
    D = len(keys)
    rb = min(root_bits, 2 * k)
    R = 1 << rb
    s = 2 * k - rb
    targets = np.arange(R + 1, dtype=np.uint64) << np.uint64(s)
    p = np.searchsorted(keys, targets[:R], side="left").astype(np.uint64)
    p = np.concatenate([p, np.array([D], np.uint64)])
    return p, s, R


def leaf_count(D, leaf_bits):
    lb = 0
    while lb < leaf_bits and (2 << lb) <= max(D, 1):
        lb += 1
    return 1 << lb


def root_ranges(p, L, D):
    R = len(p) - 1
    lo = np.minimum((p[:R] * np.uint64(L)) // np.uint64(max(D, 1)), np.uint64(L - 1)).astype(np.uint64)
    h = (p[1:] * np.uint64(L) + np.uint64(D - 1)) // np.uint64(max(D, 1)) if D else np.zeros(R, np.uint64)
    hi = np.where(h > 0, h - np.uint64(1), np.uint64(0))
    hi = np.maximum(hi, lo)
    hi = np.minimum(hi, np.uint64(L - 1))
    return lo, hi


def seg_of_keys(seg):
    """Segment id of every key from segment starts (n_segs + 1)."""
    n = np.diff(seg.astype(np.int64))
    return np.repeat(np.arange(len(n)), n)


def level_xy(level, keys, seg, s=0, lo=None, L=0):
    """FP32 inputs x and targets y of every key (device formulas)."""
    D = len(keys)
    sg = seg_of_keys(seg)
    i = np.arange(D, dtype=np.int64)
    if level == 0:
        x = (keys - (sg.astype(np.uint64) << np.uint64(s))).astype(np.float32)
        y = (i.astype(np.float64) * float(L) / float(D) - lo[sg].astype(np.float64)).astype(np.float32)
    else:
        first = keys[seg[:-1].astype(np.int64)[sg]]
        x = (keys - first).astype(np.float32)
        y = (i - seg[:-1].astype(np.int64)[sg]).astype(np.float32)
    return x, y, sg


def stats(x, sg, n_segs):
    """Mean and 1/deviation per segment (double, as the device computes
    them, before rounding to FP32)."""
    n = np.bincount(sg, minlength=n_segs).astype(np.float64)
    s1 = np.bincount(sg, weights=x.astype(np.float64), minlength=n_segs)
    s2 = np.bincount(sg, weights=x.astype(np.float64) ** 2, minlength=n_segs)
    with np.errstate(invalid="ignore", divide="ignore"):
        m = np.where(n > 0, s1 / n, 0.0)
        v = np.where(n > 0, s2 / n - m * m, 0.0)
        isd = np.where(v > 0, 1.0 / np.sqrt(v), 0.0)
    return m, isd


def residuals(x, y, sg, mu, isd, w, c):
    """FP32 residual f - y key by key (the forward kernel)."""
    mu = np.asarray(mu, np.float32)
    isd = np.asarray(isd, np.float32)
    z = ((x - mu[sg]) * isd[sg]).astype(np.float32)
    f = fma32(np.asarray(w, np.float32)[sg], z, np.asarray(c, np.float32)[sg])
    return (f - y).astype(np.float32), z


def gradients(x, y, sg, n_segs, mu, isd, w, c):
    """(2/n) sum r z, (2/n) sum r, sum r^2 per segment, in double from the
    FP32 residuals (the backward kernel)."""
    r, z = residuals(x, y, sg, mu, isd, w, c)
    r64, z64 = r.astype(np.float64), z.astype(np.float64)
    n = np.bincount(sg, minlength=n_segs).astype(np.float64)
    a1 = np.bincount(sg, weights=r64 * z64, minlength=n_segs)
    a2 = np.bincount(sg, weights=r64, minlength=n_segs)
    a3 = np.bincount(sg, weights=r64 * r64, minlength=n_segs)
    with np.errstate(invalid="ignore", divide="ignore"):
        gw = np.where(n > 0, 2.0 * a1 / n, 0.0)
        gc = np.where(n > 0, 2.0 * a2 / n, 0.0)
    return gw, gc, a3


def loss_double(x, y, sg, n_segs, mu, isd, w, c):
    """Per-segment sum of squared residuals with every operation in double
    (for finite differences)."""
    z = (x.astype(np.float64) - np.asarray(mu, np.float64)[sg]) * np.asarray(isd, np.float64)[sg]
    f = np.asarray(w, np.float64)[sg] * z + np.asarray(c, np.float64)[sg]
    r = f - y.astype(np.float64)
    return np.bincount(sg, weights=r * r, minlength=n_segs)


def least_squares(x, y, sg, n_segs, mu, isd):
    """Closed-form minimiser (w, c) of the per-segment MSE over
    z = (x - mu) * isd (double)."""
    z = (x.astype(np.float64) - np.asarray(mu, np.float64)[sg]) * np.asarray(isd, np.float64)[sg]
    y64 = y.astype(np.float64)
    n = np.bincount(sg, minlength=n_segs).astype(np.float64)
    sz = np.bincount(sg, weights=z, minlength=n_segs)
    szz = np.bincount(sg, weights=z * z, minlength=n_segs)
    sy = np.bincount(sg, weights=y64, minlength=n_segs)
    szy = np.bincount(sg, weights=z * y64, minlength=n_segs)
    det = n * szz - sz * sz
    with np.errstate(invalid="ignore", divide="ignore"):
        w = np.where(det > 0, (n * szy - sz * sy) / det, 0.0)
        c = np.where(n > 0, (sy - w * sz) / n, 0.0)
    return w, c


# ---------------------------------------------------------------- lookup arithmetic
def route(keys, root, s):
    """Leaf of every key: lo_r + clamp(floor(fma(A_r, x, B_r)), 0, hi_r - lo_r)."""
    r = (keys >> np.uint64(s)).astype(np.int64)
    P = root[r]
    x = (keys - (r.astype(np.uint64) << np.uint64(s))).astype(np.float32)
    f = np.floor(fma32(P["A"], x, P["B"])).astype(np.float32)
    span = (P["hi"] - P["lo"]).astype(np.float32)
    cl = np.where(f < 0, np.float32(0), np.where(f > span, span, f))
    return P["lo"].astype(np.int64) + cl.astype(np.int64)


def leaf_pred(keys, leaf, j):
    """q_j + floor(fma(S_j, (float)(key - base_j), I_j) + 0.5), clamped."""
    R = leaf[j]
    x = (keys - R["base"]).astype(np.int64).astype(np.float32)
    p = np.floor((fma32(R["S"], x, R["I"]) + np.float32(0.5)).astype(np.float32))
    p = np.clip(p, np.float32(-1.0e12), np.float32(1.0e12))
    return R["q"].astype(np.int64) + p.astype(np.int64)


def error_bounds(keys, leaf):
    """(elo, ehi, sum |e|) per leaf over the keys in their leaves (leaf j
    holds ranks [q_j, q_{j+1}))."""
    D = len(keys)
    L = len(leaf)
    q = np.concatenate([leaf["q"].astype(np.int64), [D]])
    j = np.repeat(np.arange(L), np.diff(q))
    e = np.arange(D, dtype=np.int64) - leaf_pred(keys, leaf, j)
    elo = np.zeros(L, np.int64)
    ehi = np.zeros(L, np.int64)
    np.maximum.at(elo, j, -e)
    np.maximum.at(ehi, j, e)
    sabs = np.bincount(j, weights=np.abs(e).astype(np.float64), minlength=L)
    return elo, ehi, sabs
