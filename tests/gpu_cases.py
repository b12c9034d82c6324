"""Input cases shared by the GPU parity tests (small enough for the oracle)."""
import random

import numpy as np

from conftest import rand_seq, rc


def edge_records():
    r = random.Random(7)
    recs = [
        "",                                   # empty record
        "ACG",                                # shorter than k
        "N" * 100,                            # all placeholder
        "A" * 300,                            # homopolymer: self loops
        "ACGT" * 60,                          # period-4 repeat
        rand_seq(r, 31),                      # exactly k (for k=31)
        rand_seq(r, 400, pn=0.03, plow=0.2),  # scattered N, lowercase
        rand_seq(r, 200).lower(),             # all lowercase
    ]
    half = rand_seq(r, 120)
    recs.append(half + rc(half))              # reverse-complement palindrome
    recs.append(rand_seq(r, 80) + "NNNNN" + rand_seq(r, 80) + "X" + rand_seq(r, 40))  # other placeholder bytes
    return [s.encode() for s in recs]


def random_genome(seed, n_rec=6, max_len=5000, pn=0.005):
    r = random.Random(seed)
    recs = []
    for _ in range(n_rec):
        s = rand_seq(r, r.randint(0, max_len), pn=pn, plow=0.05)
        if len(s) > 500 and r.random() < 0.5:  # planted repeat, maybe reverse-complemented
            a = r.randint(0, len(s) - 200)
            seg = s[a:a + 150]
            b = r.randint(0, len(s) - 200)
            s = s[:b] + (rc(seg) if r.random() < 0.5 else seg) + s[b + 150:]
        recs.append(s.encode())
    return recs
