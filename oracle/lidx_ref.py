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
  * warp_tables / warp: the learned byte-frequency CDF the models read keys
    through (fixed-point arithmetic coding, exact integer restatement);
  * route / leaf_pred / error_bounds: the lookup arithmetic (root routing,
    leaf prediction) and the recorded per-leaf error bounds;
  * encode_lines / decode_line / line_lookup: the line format and the walk
    that turns a predicted rank into the exact one.
Pinning: fma32 against exact rational arithmetic, the gradient against
finite differences of the loss and against the closed-form optimum (zero
gradient there), least_squares against numpy.linalg.lstsq, the warp against
a Fraction evaluation of its interval recurrence and its monotonicity, the
lines by a hand-built line and by finding every key from wrong starting lines
(tests/test_lidx_oracle.py).
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


# ---------------------------------------------------------------- key warp
M32 = np.uint64(0xFFFFFFFF)
WARP_MAX_GROUPS = 3


def warp_groups(key_bits):
    return min(WARP_MAX_GROUPS, key_bits // 8)


def warp_tables(keys, key_bits):
    """cum[g, s] = floor(2^32 sum_{t < s} (count_g(t) + 1) / (D + 256)) over
    the byte symbols of group g (bits [K2 - 8(g+1), K2 - 8g))."""
    G = warp_groups(key_bits)
    D = len(keys)
    cum = np.zeros((WARP_MAX_GROUPS, 256), np.uint64)
    for g in range(G):
        sym = ((keys >> np.uint64(key_bits - 8 * (g + 1))) & np.uint64(255)).astype(np.int64)
        cnt = np.bincount(sym, minlength=256).astype(object) + 1
        run = 0
        for x in range(256):
            cum[g, x] = (run << 32) // (D + 256)
            run += int(cnt[x])
    return cum[:G].astype(np.uint32)


def _f(w, c):
    """floor(w c / 2^32), w u64, c u32 (numpy uint64 arithmetic, exact)."""
    c = c.astype(np.uint64)
    return (w >> np.uint64(32)) * c + (((w & M32) * c) >> np.uint64(32))


def _mulhi64(a, b):
    a_lo, a_hi = a & M32, a >> np.uint64(32)
    b_lo, b_hi = b & M32, b >> np.uint64(32)
    ll = a_lo * b_lo
    hl = a_hi * b_lo
    lh = a_lo * b_hi
    hh = a_hi * b_hi
    cross = (ll >> np.uint64(32)) + (hl & M32) + lh
    return hh + (hl >> np.uint64(32)) + (cross >> np.uint64(32))


def warp(keys, cum, key_bits):
    """u for every key: nested intervals (lo, w) from (0, 2^64 - 1), one
    byte group at a time, then the remaining bits scaled into the last
    interval."""
    keys = np.asarray(keys, np.uint64)
    G = len(cum)
    lo = np.zeros(len(keys), np.uint64)
    w = np.full(len(keys), np.uint64(0xFFFFFFFFFFFFFFFF))
    for g in range(G):
        sym = ((keys >> np.uint64(key_bits - 8 * (g + 1))) & np.uint64(255)).astype(np.int64)
        a = _f(w, cum[g][sym])
        b = np.where(sym == 255, w, _f(w, cum[g][np.minimum(sym + 1, 255)]))
        lo = lo + a
        w = b - a
    rb = key_bits - 8 * G
    if rb > 0:
        rest = (keys & np.uint64((1 << rb) - 1)) << np.uint64(64 - rb)
        lo = lo + _mulhi64(w, rest)
    return lo


# ---------------------------------------------------------------- segments
def root_pieces(keys, root_bits):
    """Piece starts p[0..R] (first rank with warped key >= r << s, s = 64 -
    root_bits) over the warped keys; also returns s, R."""
    D = len(keys)
    rb = max(1, min(root_bits, 20))
    R = 1 << rb
    s = 64 - rb
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
    """lo_r = floor(p_r L / D) (the root's training offset, r < R) and the
    first leaf of every piece, leaf0_r = lo_r + r (r <= R): piece r owns
    leaves [leaf0_r, leaf0_{r+1}), at least one, and no leaf holds keys of two
    pieces. L + R leaves in all."""
    R = len(p) - 1
    lo = (p.astype(np.uint64) * np.uint64(L)) // np.uint64(max(D, 1)) if D else np.zeros(R + 1, np.uint64)
    leaf0 = lo + np.arange(R + 1, dtype=np.uint64)
    return lo[:R], leaf0


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
ROOT_DTYPE = np.dtype([("A", "<f4"), ("B", "<f4"), ("leaf0", "<u4"), ("p", "<u4")])
LEAF_DTYPE = np.dtype([("S", "<f4"), ("I", "<f4")])


def piece_x(keys, s):
    """x = key - (r << s) in FP32 from its top 32 bits (s > 32):
    (float)(u32)(key >> (s - 32)) * 2^(s - 32)."""
    top = ((keys >> np.uint64(s - 32)) & np.uint64(0xFFFFFFFF)).astype(np.uint32)
    return (top.astype(np.float32) * np.float32(2.0 ** (s - 32))).astype(np.float32)


def route(keys, root, s):
    """Leaf of every key: leaf0_r + clamp(floor(fma(A_r, x, B_r)), 0, n_r - 1)
    with x = piece_x(key) and n_r = leaf0_{r+1} - leaf0_r."""
    r = (keys >> np.uint64(s)).astype(np.int64)
    P = root[r]
    n = root["leaf0"][r + 1].astype(np.int64) - P["leaf0"].astype(np.int64)
    x = piece_x(keys, s)
    f = np.floor(fma32(P["A"], x, P["B"])).astype(np.float64)
    cl = np.clip(f, 0, np.maximum(n - 1, 0)).astype(np.int64)
    return P["leaf0"].astype(np.int64) + cl


def is_fallback(leaf):
    """Leaves whose slope carries the sign bit resolve through the BBHash."""
    return (leaf["S"].view(np.uint32) >> np.uint32(31)).astype(bool)


def leaf_pred(keys, root, leaf, s, j):
    """p_r + floor(fma(|S_j|, piece_x(key), I_j) + 0.5), the local part
    clamped to +-2^30: the rank the leaf predicts (piece coordinates)."""
    r = (keys >> np.uint64(s)).astype(np.int64)
    x = piece_x(keys, s)
    S = np.abs(leaf["S"][j]).astype(np.float32)
    f = (fma32(S, x, leaf["I"][j]) + np.float32(0.5)).astype(np.float32)
    f = np.clip(np.floor(f), np.float32(-1073741824.0), np.float32(1073741824.0))
    return root["p"][r].astype(np.int64) + f.astype(np.int64)


def leaf_starts(keys, root, s, n_leaves):
    """q[j] = first rank routed to a leaf >= j, j in [0, n_leaves]."""
    j = route(keys, root, s) if len(keys) else np.zeros(0, np.int64)
    return np.searchsorted(j, np.arange(n_leaves + 1), side="left").astype(np.int64)


def error_bounds(keys, root, leaf, s):
    """(elo, ehi, sum |e|) per leaf over the keys routed to it, with the
    lookup's own arithmetic: rank - pred in [-elo, ehi]."""
    D = len(keys)
    L = len(leaf)
    j = route(keys, root, s) if D else np.zeros(0, np.int64)
    e = np.arange(D, dtype=np.int64) - leaf_pred(keys, root, leaf, s, j)
    elo = np.zeros(L, np.int64)
    ehi = np.zeros(L, np.int64)
    np.maximum.at(elo, j, -e)
    np.maximum.at(ehi, j, e)
    sabs = np.bincount(j, weights=np.abs(e).astype(np.float64), minlength=L)
    return elo, ehi, sabs


# ---------------------------------------------------------------- lines
# A line holds C consecutive keys by rank (line i: ranks [iC, iC + n_i)) plus
# the next line's first key (the sentinel, rank iC + C) when there is one, in
# line_bytes bytes read as a little-endian bit stream of u64 words:
#   bits  0..61  k0, the key of rank iC;  bit 62: overflow (no payload);
#   bit 63: code width w (0: 3 bits, 1: 4 bits)
#   bits 64..69  t, the cell shift;        bits 70..79: ulen
#   bits 80..87  the sentinel's prefix: the top 8 bits of its offset in its
#         cell (sentinel_prefix; 0 without a sentinel)
#   unary [88, 88 + ulen): for cells c = 0 .. (k_last - k0) >> t, one 1 per
#         key with (key - k0) >> t == c, then a 0
#   codes after the unary: for every key j >= 1 in the cell of key j-1, in
#         order, w bits min(t - bitlen((k_j - k0) ^ (k_{j-1} - k0)), 2^w - 1):
#         the highest bit where the two differ is t - 1 - code
# For each w, t_w minimises m + cells + w collisions over [tb - 2, tb + 3]
# (first minimum), tb = max(0, bitlen(span) - bitlen(m)). The line takes
# 4-bit codes when they fit 8 line_bytes - 88 bits and 3-bit codes would not
# fit or would hold a code of 7 (some pair closer than t_3 - 7 bits); else
# 3-bit codes when they fit; else it is an overflow line.
HDR_BITS = 88
SP_BITS = 8


def sentinel_prefix(rel, t):
    """Top SP_BITS bits of rel's offset in its cell (rel mod 2^t), read as a
    t-bit number (exact, shifted up, when t < SP_BITS): of two values in one
    cell, the one with the larger prefix is the larger."""
    off = rel & ((1 << t) - 1)
    return off >> (t - SP_BITS) if t >= SP_BITS else off << (SP_BITS - t)


def default_line_keys(line_bytes):
    return {64: 108, 128: 256}[line_bytes]


def n_lines(D, C):
    return (D + C - 1) // C


def _line_keys(keys, i, C):
    D = len(keys)
    a = i * C
    n = min(C, D - a)
    m = n + (1 if a + n < D else 0)
    return [int(x) for x in keys[a:a + m]], n, m


def line_choice(ks, line_bytes):
    """(t, bits, collisions, code width) chosen for the keys of one line
    (bits above the payload: an overflow line)."""
    m = len(ks)
    k0 = ks[0]
    rel = [x - k0 for x in ks]
    span = rel[-1]
    bl = [(rel[j] ^ rel[j - 1]).bit_length() for j in range(1, m)]
    min_bl = min(bl) if bl else 64
    tb = max(0, span.bit_length() - m.bit_length())
    best = {3: None, 4: None}
    for t in range(max(0, tb - 2), min(61, tb + 3) + 1):
        coll = sum(1 for b in bl if b <= t)
        for w in (3, 4):
            bits = m + (span >> t) + 1 + w * coll
            if best[w] is None or bits < best[w][1]:
                best[w] = (t, bits, coll)
    payload = 8 * line_bytes - HDR_BITS
    t3, b3, _ = best[3]
    if best[4][1] <= payload and (b3 > payload or (min_bl <= t3 and t3 - min_bl >= 7)):
        return best[4] + (4,)
    return best[3] + (3,)


def encode_line(keys, i, C, line_bytes):
    """The line_bytes bytes of line i and (bits used, overflow, keys of runs
    that resolve by search)."""
    ks, n, m = _line_keys(keys, i, C)
    nw = line_bytes // 8
    w = [0] * nw

    def put(pos, val, width):
        for b in range(width):
            if (val >> b) & 1:
                w[(pos + b) >> 6] |= 1 << ((pos + b) & 63)

    k0 = ks[0]
    put(0, k0, 62)
    t, bits, coll, cw = line_choice(ks, line_bytes)
    if bits > 8 * line_bytes - HDR_BITS:
        put(62, 1, 1)
        return np.array(w, np.uint64).view(np.uint8), (0, True, n)
    if cw == 4:
        put(63, 1, 1)
    rel = [x - k0 for x in ks]
    cells = (rel[-1] >> t) + 1
    ulen = m + cells
    put(64, t, 6)
    put(70, ulen, 10)
    if m > n:
        put(80, sentinel_prefix(rel[-1], t), SP_BITS)
    pos, prev = HDR_BITS, 0
    for j in range(m):
        c = rel[j] >> t
        pos += c - prev
        prev = c
        put(pos, 1, 1)
        pos += 1
    pos += 1
    assert pos == HDR_BITS + ulen
    # codes, and the keys whose run holds a search code or too many codes
    csearch = (1 << cw) - 1
    run_len, run_bad, n_search = 1, False, 0
    for j in range(1, m + 1):
        same = j < m and (rel[j] >> t) == (rel[j - 1] >> t)
        if same:
            code = min(t - (rel[j] ^ rel[j - 1]).bit_length(), csearch)
            put(pos, code, cw)
            pos += cw
            run_len += 1
            run_bad |= code == csearch
        else:
            if run_len > 1 and (run_bad or run_len - 1 > 64 // cw):
                n_search += run_len
            run_len, run_bad = 1, False
    return np.array(w, np.uint64).view(np.uint8), (bits, False, n_search)


def encode_lines(keys, C, line_bytes):
    """All lines (n_lines x line_bytes bytes) and per-line statistics."""
    nl = n_lines(len(keys), C)
    out = np.zeros((nl, line_bytes), np.uint8)
    st = []
    for i in range(nl):
        out[i], s = encode_line(keys, i, C, line_bytes)
        st.append(s)
    return out, st


def decode_line(line, q, i, C, n_total_lines):
    """One step of the walk for query q at line i: ("found", rank),
    ("left",), ("right",), ("absent",) or ("search", a, b, right_ok): binary
    search ranks [a, b] of the sorted keys; right_ok: a q above them all goes
    right. In the run that holds the sentinel, the prefixes of q and the
    sentinel decide: a larger one goes right, a smaller one leaves the run
    without the sentinel, and equal ones search the run."""
    w = [int(x) for x in np.asarray(line).view(np.uint64)]

    def bit(b):
        return (w[b >> 6] >> (b & 63)) & 1

    def get(pos, width):
        return sum(bit(pos + b) << b for b in range(width))

    has_sent = i + 1 < n_total_lines
    k0 = w[0] & ((1 << 62) - 1)
    if (w[0] >> 62) & 1:
        if q < k0:
            return ("left",)
        return ("search", i * C, i * C + C if has_sent else None, has_sent)
    if q < k0:
        return ("left",)
    t = get(64, 6)
    ulen = get(70, 10)
    cw = 4 if (w[0] >> 63) & 1 else 3
    rel = q - k0
    cq = rel >> t
    zeros = [b for b in range(HDR_BITS, HDR_BITS + ulen) if not bit(b)]
    if cq >= len(zeros):
        return ("right",)
    e_b = zeros[cq]
    e_a = zeros[cq - 1] if cq > 0 else HDR_BITS - 1
    s = e_b - e_a - 1
    j0 = (e_a + 1 - HDR_BITS) - cq
    if s == 0:
        return ("absent",)
    last = j0 + s - 1
    if has_sent and last == C:
        qp, sp = sentinel_prefix(rel, t), get(80, SP_BITS)
        if qp > sp:
            return ("right",)
        if qp == sp:
            return ("search", i * C + j0, i * C + last, True)
        s, last = s - 1, last - 1
        if s == 0:
            return ("absent",)
    if s == 1:
        return ("found", i * C + j0)
    if s - 1 > 64 // cw:
        return ("search", i * C + j0, i * C + last, False)
    ne = sum(1 for b in range(HDR_BITS + 1, e_a + 1) if not bit(b) and bit(b - 1))
    m0 = j0 - ne
    cpos = HDR_BITS + ulen + cw * m0
    codes = [get(cpos + cw * x, cw) for x in range(s - 1)]
    if (1 << cw) - 1 in codes:
        return ("search", i * C + j0, i * C + last, False)
    lo, hi = 0, s - 1
    while lo < hi:
        b = min(range(lo + 1, hi + 1), key=lambda x: codes[x - 1])
        d = t - 1 - codes[b - 1]
        if (rel >> d) & 1:
            lo = b
        else:
            hi = b - 1
    return ("found", i * C + j0 + lo)


def line_lookup(keys, lines, C, q, start, max_steps=64):
    """Rank of q by the walk from line `start` (None when absent)."""
    nl = len(lines)
    if nl == 0:
        return None
    i = min(max(start, 0), nl - 1)
    for _ in range(max_steps):
        r = decode_line(lines[i], q, i, C, nl)
        if r[0] == "found":
            return r[1]
        if r[0] == "absent":
            return None
        if r[0] == "left":
            if i == 0:
                return None
            i -= 1
            continue
        if r[0] == "right":
            if i == nl - 1:
                return None
            i += 1
            continue
        a, b, right_ok = r[1], r[2], r[3]
        if b is None:
            b = len(keys) - 1
        x = int(np.searchsorted(keys[a:b + 1], np.uint64(q)))
        if x <= b - a and int(keys[a + x]) == q:
            return a + x
        if right_ok and x == b - a + 1 and i < nl - 1:
            i += 1
            continue
        return None
    return None
