"""Seeded synthetic inputs for the benchmark configurations.

Each generator returns (ascii uint8 array, rec_off uint64[n+1], names) with
records concatenated, ready for Cf1g.load_records(ascii=..., rec_off=...).

  genome_5m     : one 5 Mbp record, i.i.d. uniform ACGT (seed 1), plus 200
                  planted 2 kbp exact repeats (100 reverse-complemented) at
                  uniform positions.
  transcriptome : 100k records, lognormal lengths (median 1200, sigma 0.9,
                  clipped 50..100000) built from a pool of 400k random 150 bp
                  exons (seed 2), 3..15 exons per record, concatenated then
                  tiled/truncated to the drawn length; 2% of records shorter
                  than k (uniform 1..k-1).
  genome_300m   : 24 records totalling 300 Mbp, Zipf-like lengths (largest
                  40 Mbp), GC 41% drifting per 10 kbp, 5% of sequence from 8
                  repeat families (300 bp..6 kbp, 10-20% point divergence),
                  poly-N runs of 1..50 kbp every ~5 Mbp (seed 3).
All sizes scale with `scale` for small test versions of the same shapes.
"""
import numpy as np

CODE2ASCII = np.frombuffer(b"ACGTN", dtype=np.uint8)


def _rc_codes(c):
    out = (3 - c[::-1]).astype(np.uint8)
    out[c[::-1] == 4] = 4
    return out


def _finish(codes_list, names):
    off = np.zeros(len(codes_list) + 1, np.uint64)
    off[1:] = np.cumsum([len(c) for c in codes_list], dtype=np.uint64)
    ascii_ = np.empty(int(off[-1]), np.uint8)
    for i, c in enumerate(codes_list):
        ascii_[int(off[i]):int(off[i + 1])] = CODE2ASCII[c]
    return ascii_, off, names


def genome_5m(scale=1.0, seed=1, n_repeats=200, repeat_len=2000):
    rng = np.random.default_rng(seed)
    n = int(5_000_000 * scale)
    g = rng.integers(0, 4, n, dtype=np.uint8)
    nr = max(1, int(n_repeats * min(1.0, scale * 4))) if scale < 1 else n_repeats
    rl = min(repeat_len, max(n // 10, 1))
    for i in range(nr):
        src = int(rng.integers(0, n - rl))
        dst = int(rng.integers(0, n - rl))
        seg = g[src:src + rl].copy()
        if i % 2 == 1:
            seg = _rc_codes(seg)
        g[dst:dst + rl] = seg
    return _finish([g], ["chr1"])


def transcriptome(scale=1.0, seed=2, k=31, n_records=100_000, n_exons=400_000, exon_len=150):
    rng = np.random.default_rng(seed)
    nrec = max(1, int(n_records * scale))
    nex = max(16, int(n_exons * scale))
    pool = rng.integers(0, 4, (nex, exon_len), dtype=np.uint8)
    lens = np.exp(rng.normal(np.log(1200.0), 0.9, nrec))
    lens = np.clip(lens, 50, 100_000).astype(np.int64)
    short = rng.random(nrec) < 0.02
    lens[short] = rng.integers(1, max(2, k), int(short.sum()))
    nper = rng.integers(3, 16, nrec)
    recs = []
    for r in range(nrec):
        ids = rng.integers(0, nex, int(nper[r]))
        cat = pool[ids].ravel()
        recs.append(np.resize(cat, int(lens[r])))
    return _finish(recs, [f"tx{r:06d}" for r in range(nrec)])


def _zipf_lengths(total, n, largest):
    lo, hi = 0.0, 3.0
    for _ in range(100):
        s = (lo + hi) / 2
        w = np.arange(1, n + 1, dtype=np.float64) ** -s
        if w[0] / w.sum() * total > largest:
            hi = s
        else:
            lo = s
    w = np.arange(1, n + 1, dtype=np.float64) ** -((lo + hi) / 2)
    L = np.floor(w / w.sum() * total).astype(np.int64)
    L[-1] += total - int(L.sum())
    return L


def genome_300m(scale=1.0, seed=3, n_records=24):
    rng = np.random.default_rng(seed)
    total = int(300_000_000 * scale)
    lens = _zipf_lengths(total, n_records, 40_000_000 * scale)
    fam = []
    for _ in range(8):
        fl = int(rng.integers(300, 6001))
        fam.append(rng.integers(0, 4, fl, dtype=np.uint8))
    recs = []
    for L in lens:
        L = int(L)
        nb = (L + 9999) // 10000
        gc = np.empty(nb, np.float64)
        g = 0.41
        noise = rng.normal(0.0, 0.02, nb)
        for b in range(nb):
            g = 0.41 + 0.9 * (g - 0.41) + noise[b]
            gc[b] = min(max(g, 0.25), 0.6)
        gcpos = np.repeat(gc.astype(np.float32), 10000)[:L]
        u = rng.random(L, dtype=np.float32)
        v = rng.integers(0, 2, L, dtype=np.uint8)
        is_gc = u < gcpos
        codes = np.where(is_gc, 1 + v, np.where(v == 1, 3, 0)).astype(np.uint8)
        del u, v, is_gc, gcpos
        recs.append(codes)
    # interspersed repeats: 5% of the sequence, 10-20% point divergence
    target = int(0.05 * total)
    weights = lens / lens.sum()
    placed = 0
    while placed < target:
        f = fam[int(rng.integers(0, 8))]
        d = float(rng.uniform(0.10, 0.20))
        cp = f.copy()
        mut = rng.random(len(cp)) < d
        cp[mut] = (cp[mut] + rng.integers(1, 4, int(mut.sum()), dtype=np.uint8)) % 4
        if rng.random() < 0.5:
            cp = _rc_codes(cp)
        r = int(rng.choice(len(recs), p=weights))
        if len(recs[r]) <= len(cp):
            continue
        pos = int(rng.integers(0, len(recs[r]) - len(cp)))
        recs[r][pos:pos + len(cp)] = cp
        placed += len(cp)
    # poly-N runs every ~5 Mbp
    for c in recs:
        pos = int(rng.integers(1_000_000, 5_000_000)) if len(c) > 5_000_000 else len(c)
        while pos < len(c):
            nl = int(rng.integers(1000, 50_001))
            c[pos:pos + nl] = 4
            pos += nl + int(rng.integers(4_000_000, 6_000_000))
    return _finish(recs, [f"chr{i + 1}" for i in range(len(recs))])


def write_fasta(path, ascii_, rec_off, names, width=80):
    with open(path, "wb") as f:
        for i, nm in enumerate(names):
            f.write(b">" + nm.encode() + b"\n")
            s = ascii_[int(rec_off[i]):int(rec_off[i + 1])].tobytes()
            for j in range(0, len(s), width):
                f.write(s[j:j + width] + b"\n")


def records_list(ascii_, rec_off):
    return [ascii_[int(rec_off[i]):int(rec_off[i + 1])].tobytes() for i in range(len(rec_off) - 1)]
