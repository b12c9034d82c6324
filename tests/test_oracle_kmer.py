"""k-mer encoding, reverse complement, canonical form, rolling and 2-bit
decode (src/kmer.rs) against string-level brute force for every K."""
import ctypes as C
import random

import numpy as np
import pytest

from conftest import rc
from oracle import lib

KS = list(range(1, 64, 2)) + [4, 32]


def enc(s):
    v = 0
    for c in s:
        v = (v << 2) | "ACGT".index(c.upper())
    return v


def kops(k, s, op):
    out = (C.c_uint64 * 4)()
    lab = C.create_string_buffer(2 * k + 4)
    assert lib().ref_kmer_ops(k, s.encode(), op, out, lab) == 0
    return (out[1] << 64) | out[0], lab.raw[:k].decode(), lab.raw[k + 1:2 * k + 1].decode(), list(out)


@pytest.mark.parametrize("k", [k for k in KS if k != 32 or True])
def test_encode_rc_canonical_roll(k):
    if k == 32:
        pytest.skip("k=32 has no dispatch in the reference (odd k only)")
    r = random.Random(k)
    for _ in range(60):
        s = "".join(r.choice("ACGTacgt") for _ in range(k + 1))
        x, lab, _, _ = kops(k, s, 0)
        assert x == enc(s[:k]) and lab == s[:k].upper()
        y, rl, _, _ = kops(k, s, 1)
        assert rl == rc(s[:k]) and y == enc(rc(s[:k]))
        c, cl, _, _ = kops(k, s, 2)
        assert c == min(enc(s[:k]), enc(rc(s[:k])))
        n, nl, nrl, _ = kops(k, s, 3)
        assert nl == s[1:k + 1].upper() and nrl == rc(s[1:k + 1])
        _, _, _, fb = kops(k, s, 4)
        assert "ACGT"[fb[2]] == s[0].upper() and "ACGT"[fb[3]] == s[k - 1].upper()


@pytest.mark.parametrize("k", [1, 3, 31, 33, 61, 63])
def test_from_packed_2bit_all_offsets(k):
    # Decoding at every base offset, including K=63 offsets = 2, 3 (mod 4),
    # where the reference's u128 accumulator is too narrow (src/kmer.rs:148):
    # the oracle decodes by the format's definition.
    r = random.Random(100 + k)
    s = "".join(r.choice("ACGT") for _ in range(k + 40))
    packed = np.zeros((len(s) + 3) // 4, np.uint8)
    for j, ch in enumerate(s):
        packed[j // 4] |= "ACGT".index(ch) << (6 - 2 * (j % 4))
    out = (C.c_uint64 * 2)()
    for off in range(0, 40):
        lib().ref_kmer_from_packed(k, packed.ctypes.data_as(C.POINTER(C.c_uint8)), off, out)
        assert (out[1] << 64) | out[0] == enc(s[off:off + k]), off
