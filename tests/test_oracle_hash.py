"""XXH3 bucket hash (src/minimizer.rs:18-21) against python xxhash vectors."""
import json
import os

from oracle import lib

G = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "xxh3_vectors.json")))


def test_xxh3_golden_vectors():
    L = lib()
    assert len(G["cases"]) > 4000
    for v, h, b in G["cases"]:
        assert L.ref_xxh3_u64(int(v)) == int(h), v
        assert L.ref_bucket_hash(int(v)) == b, v


def test_bucket_range():
    L = lib()
    assert all(0 <= L.ref_bucket_hash(v) < 65536 for v in range(0, 100000, 7))
