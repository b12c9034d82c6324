"""The reference's own unit tests (known answers), run against the oracle."""
import ctypes as C
import json
import os

import numpy as np

from oracle import lib

KAT = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_kat.json")))


def kops(k, s, op):
    out = (C.c_uint64 * 4)()
    lab = C.create_string_buffer(2 * k + 4)
    assert lib().ref_kmer_ops(k, s.encode(), op, out, lab) == 0
    return list(out), lab.raw[:k].decode(), lab.raw[k + 1:2 * k + 1].decode()


def test_dna_tables():  # src/dna.rs:111-131
    cls = np.zeros(32, np.uint8)
    for a, b in KAT["dna"]["complement"]:
        # complement of a base = 3 - code; checked through rc of a 1-mer
        assert kops(1, a, 1)[1] == b
    for ch, code in KAT["dna"]["from_ascii"]:
        if code is not None:
            assert kops(1, ch, 0)[0][0] == code


def test_placeholders():
    # is_placeholder: a 1-mer run over a placeholder byte yields no k-mer (D=0)
    from oracle import RefPipeline
    for ch, ph in KAT["dna"]["is_placeholder"]:
        p = RefPipeline(1, num_bins=1)
        p.load([ch.encode()])
        assert p.run_p1_p4() == (0 if ph else 1)


def test_kmer_label_roundtrip():  # kmer.rs:509-515
    for k, s in KAT["kmer"]["label_roundtrip"]:
        assert kops(k, s, 0)[1] == s


def test_kmer_rc_canonical_front_back_roll():  # kmer.rs:517-561
    for k, s, r in KAT["kmer"]["reverse_complement"]:
        assert kops(k, s, 1)[1] == r
    for k, s, c in KAT["kmer"]["canonical"]:
        assert kops(k, s, 2)[1] == c
    for k, s, f, b in KAT["kmer"]["front_back"]:
        o = kops(k, s, 4)[0]
        assert "ACGT"[o[2]] == f and "ACGT"[o[3]] == b
    for k, s, nk, nrc in KAT["kmer"]["roll"]:
        _, lab, rlab = kops(k, s, 3)
        assert lab == nk and rlab == nrc


def test_kmer_ordering_and_k33():  # kmer.rs:572-591
    for k, seq in KAT["kmer"]["ordering"]:
        vals = [kops(k, s, 0)[0] for s in seq]
        keys = [(v[1], v[0]) for v in vals]
        assert keys == sorted(keys) and len(set(keys)) == len(keys)
    for k, s, lab in KAT["kmer"]["k33"]:
        o, l, _ = kops(k, s, 0)
        assert l == lab
        r, _, _ = kops(k, s, 1)
        # rc(rc(x)) == x through the label of rc
        rl = kops(k, s, 1)[1]
        assert kops(k, rl, 1)[1] == lab


def test_state_codes():  # src/state.rs:304-348
    L = lib()
    st = KAT["state"]
    assert L.ref_state_from_vertex(0, st["siso_CG"]["front"], st["siso_CG"]["back"], 1, 0) == st["siso_CG"]["code"]
    assert L.ref_state_from_vertex(3, 0, 0, 1, 0) == st["mimo"]["code"]
    assert L.ref_state_from_vertex(1, 0, st["miso_T"]["back"], 1, 0) == st["miso_T"]["code"]
    assert L.ref_state_from_vertex(2, st["simo_G"]["front"], 0, 1, 0) == st["simo_G"]["code"]
    cls, outp, dead, fr, bk = (np.zeros(32, np.uint8) for _ in range(5))
    p = lambda a: a.ctypes.data_as(C.POINTER(C.c_uint8))
    L.ref_state_tables(p(cls), p(outp), p(dead), p(fr), p(bk))
    assert outp[22] == st["outputted_of_22"]
    assert dead[3] == 1
    assert cls[st["siso_CG"]["code"]] == st["siso_CG"]["class"]
    assert cls[7] == st["miso_T"]["class"] and cls[10] == st["simo_G"]["class"]
    assert fr[22] == 1 and bk[22] == 2


def test_state_vector():  # src/state_vector.rs:81-113
    L = lib()
    sv = KAT["state_vector"]
    h = L.ref_sv_new(sv["basic"]["len"])
    for op in sv["basic"]["ops"]:
        if op[0] == "get":
            assert L.ref_sv_get(h, op[1]) == op[2]
        else:
            assert bool(L.ref_sv_cas(h, op[1], op[2], op[3])) == op[4]
    L.ref_sv_free(h)
    h = L.ref_sv_new(sv["all_positions"]["len"])
    for i, v in enumerate(sv["all_positions"]["values"]):
        assert L.ref_sv_cas(h, i, 0, v) == 1 and L.ref_sv_get(h, i) == v
    # 12 states fill one word, 5 bits each, no straddling
    w = L.ref_sv_word(h, 0)
    assert all(((w >> (5 * i)) & 31) == v for i, v in enumerate(sv["all_positions"]["values"]))
    assert w >> 60 == 0
    L.ref_sv_free(h)
