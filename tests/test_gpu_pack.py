"""The device packing of ASCII records into the flat 2-bit stream (run by P1
after cf1g_load_records) against a numpy statement of the layout: position
0 and the position after each record are separators, record r's bases start
at off[r] + r + 1, A/C/G/T of either case are codes 0-3 and every other byte
is a placeholder (code 0, N bit set); word w holds positions 64w .. 64w+63,
MSB-first, and positions past the stream are placeholders. Records of every
length from 0 to past 64 bases start at every byte alignment, so words
inside one record and words that straddle separators both occur."""
import random

import numpy as np
import pytest

import cf1g

pytestmark = pytest.mark.gpu

CODE = np.full(256, 4, np.uint8)
for ch, v in ((b"A", 0), (b"C", 1), (b"G", 2), (b"T", 3)):
    CODE[ch[0]] = v
    CODE[ch.lower()[0]] = v


def expected_stream(recs):
    n_ascii = sum(len(s) for s in recs)
    n_pos = n_ascii + len(recs) + 1
    nw = (n_pos + 63) // 64
    codes = np.full(nw * 64, 4, np.uint8)
    g = 1
    for s in recs:
        codes[g:g + len(s)] = CODE[np.frombuffer(s, np.uint8)] if s else codes[g:g]
        g += len(s) + 1
    c = codes.reshape(nw, 64)
    shifts = np.arange(62, -1, -2, dtype=np.uint64)
    b = (c & 3).astype(np.uint64)
    hi = np.bitwise_or.reduce(b[:, :32] << shifts, axis=1)
    lo = np.bitwise_or.reduce(b[:, 32:] << shifts, axis=1)
    seq = np.empty(2 * nw, np.uint64)
    seq[0::2], seq[1::2] = hi, lo
    nbit = (c >> 2).astype(np.uint64)
    nmask = np.bitwise_or.reduce(nbit << np.arange(63, -1, -1, dtype=np.uint64), axis=1)
    return seq, nmask


def records(seed):
    r = random.Random(seed)
    recs = []
    for L in list(range(0, 140)) + [r.randint(200, 5000) for _ in range(30)]:
        kind = r.random()
        if kind < 0.6:  # genome-like: mostly ACGT, some lowercase and N
            s = bytes(r.choice(b"ACGTACGTACGTacgtN") for _ in range(L))
        elif kind < 0.8:  # every byte value
            s = bytes(r.randrange(256) for _ in range(L))
        else:  # letters next to ACGT in the byte table, both cases
            s = bytes(r.choice(b"ABCDEFGHSTUVWacgtbdhsu@[`{\x80\xc1\xc3\xc7\xd4") for _ in range(L))
        recs.append(s)
    r.shuffle(recs)
    return recs


@pytest.mark.parametrize("seed", [1, 2])
def test_device_packing_matches_layout(seed):
    recs = records(seed)
    g = cf1g.Cf1g(k=31, num_bins=8)
    g.load_records(recs)
    g.count_minimizer_histogram(want=False)  # packs the stream
    seq, nmask = g.stream_export()
    want_seq, want_nmask = expected_stream(recs)
    np.testing.assert_array_equal(nmask, want_nmask)
    np.testing.assert_array_equal(seq, want_seq)
    g.close()


def test_device_packing_long_record_offsets():
    # one long record after a short one of each length 0..3, so that the long
    # record's bytes start at each alignment mod 4; words inside it take the
    # four-bytes-at-a-time path
    r = random.Random(5)
    long_rec = bytes(r.choice(b"ACGTacgtN") for _ in range(20000))
    for lead in range(4):
        recs = [b"A" * lead, long_rec, b"GATTACA"]
        g = cf1g.Cf1g(k=31, num_bins=4)
        g.load_records(recs)
        g.count_minimizer_histogram(want=False)
        seq, nmask = g.stream_export()
        want_seq, want_nmask = expected_stream(recs)
        np.testing.assert_array_equal(nmask, want_nmask, err_msg=f"lead {lead}")
        np.testing.assert_array_equal(seq, want_seq, err_msg=f"lead {lead}")
        g.close()
