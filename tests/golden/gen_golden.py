"""Writes the golden fixtures under tests/golden/.

xxh3_vectors.json: xxh3_64 (seed 0) of the 8 little-endian bytes of u64
values and the 16-bit bucket (>> 48) that src/minimizer.rs:18-21 derives,
computed with the python `xxhash` package (3.8.1), the library routine the
reference's xxhash-rust 0.8.15 implements.

reference_kat.json: the inputs and expected outputs of the reference's own
unit tests (src/dna.rs:107-134, src/kmer.rs:505-592, src/state.rs:300-349,
src/state_vector.rs:77-114), as data.
"""
import json
import os
import random

import xxhash

HERE = os.path.dirname(os.path.abspath(__file__))


def xxh3_vectors():
    vals = [0, 1, 2, 3, 4, 0xFF, 0xFFFF, 0xFFFFFFFF, 2**64 - 1, 2**63, 2**32, 0x0123456789ABCDEF]
    vals += [1 << i for i in range(64)]
    vals += [(1 << i) - 1 for i in range(1, 64)]
    vals += list(range(5, 1029))  # small canonical m-mer values
    r = random.Random(20240601)
    vals += [r.getrandbits(30) for _ in range(1500)]  # m <= 15 minimizer values
    vals += [r.getrandbits(64) for _ in range(1500)]
    out = []
    for v in vals:
        h = xxhash.xxh3_64_intdigest(v.to_bytes(8, "little"))
        out.append([str(v), str(h), h >> 48])
    return {"library": "python xxhash " + xxhash.VERSION, "cases": out}


def reference_kat():
    return {
        "dna": {
            "complement": [["A", "T"], ["C", "G"], ["G", "C"], ["T", "A"]],
            "from_ascii": [["A", 0], ["a", 0], ["N", None]],
            "is_placeholder": [["A", False], ["a", False], ["N", True], ["n", True], ["X", True]],
        },
        "kmer": {
            "label_roundtrip": [[31, "ACGTACGTACGTACGTACGTACGTACGTACG"]],
            "reverse_complement": [[4, "ACGT", "ACGT"], [4, "AAAC", "GTTT"]],
            "canonical": [[4, "AAAC", "AAAC"]],
            "front_back": [[4, "ACGT", "A", "T"]],
            "roll": [[4, "ACGTG", "CGTG", "CACG"]],
            "ordering": [[4, ["AAAA", "AAAC", "TTTT"]]],
            "k33": [[33, "ACGTACGTACGTACGTACGTACGTACGTACGTACG", "ACGTACGTACGTACGTACGTACGTACGTACGTA"]],
        },
        "state": {
            "siso_CG": {"front": 1, "back": 2, "code": 22, "class": 0},
            "mimo": {"code": 3, "dead_end": True},
            "outputted_of_22": 12,
            "miso_T": {"back": 3, "code": 7, "class": 1},
            "simo_G": {"front": 2, "code": 10, "class": 2},
        },
        "state_vector": {
            "basic": {"len": 100, "ops": [["get", 0, 0], ["get", 99, 0], ["cas", 0, 0, 3, True], ["get", 0, 3],
                                          ["cas", 0, 0, 5, False], ["get", 0, 3], ["cas", 11, 0, 16, True],
                                          ["get", 11, 16], ["cas", 12, 0, 31, True], ["get", 12, 31]]},
            "all_positions": {"len": 12, "values": [max(i % 32, 3) for i in range(12)]},
        },
    }


if __name__ == "__main__":
    with open(os.path.join(HERE, "xxh3_vectors.json"), "w") as f:
        json.dump(xxh3_vectors(), f)
    with open(os.path.join(HERE, "reference_kat.json"), "w") as f:
        json.dump(reference_kat(), f, indent=1)
    print("wrote golden fixtures")
