"""Host ingest (cf1g_load_fasta): FASTA / gzip parsed on the host, the 2-bit
stream packed on the host. The records must equal a plain parse of the same
files, and the packed stream and every phase product must equal those of the
same records given as ASCII to cf1g_load_records (packed on the device)."""
import gzip
import os
import random
import tempfile

import numpy as np
import pytest

import cf1g
from cf1g import synth
from conftest import rand_seq

pytestmark = pytest.mark.gpu


def parse(text):
    """Reference reading of FASTA text: '>' lines start records (id up to the
    first whitespace), other lines are sequence without their line breaks."""
    recs, names, cur = [], [], None
    for line in text.split(b"\n"):
        if line.startswith(b">"):
            if cur is not None:
                recs.append(b"".join(cur))
            names.append(line[1:].split()[0].decode() if line[1:].split() else "")
            cur = []
        elif cur is not None:
            cur.append(line.replace(b"\r", b""))
    if cur is not None:
        recs.append(b"".join(cur))
    return recs, names


def odd_fasta(seed):
    r = random.Random(seed)
    parts = [b"junk before the first header\n"]
    for i in range(12):
        seq = rand_seq(r, r.choice([0, 5, 31, 200, 3000]), pn=0.02, plow=0.1, alphabet="ACGTRY").encode()
        width = r.choice([7, 60, 80, 10 ** 6])
        lines = [seq[j:j + width] for j in range(0, len(seq), width)] or [b""]
        eol = b"\r\n" if i % 3 == 0 else b"\n"
        parts.append(b">rec%d%s" % (i, b" some description\there" if i % 2 else b"") + eol + eol.join(lines) + eol)
    parts.append(b">last_no_newline\nACGTACGTNNACGT")
    return b"".join(parts)


def products(g):
    g.run_p1_p4()
    return g.mphf_export(), g.states(), g.annotate()


def check_same(a, b):
    for x, y, what in zip(a, b, ("MPHF", "states", "annotation")):
        np.testing.assert_array_equal(x, y, err_msg=what)


def test_fasta_and_gzip_records_and_products():
    with tempfile.TemporaryDirectory() as td:
        t1, t2 = odd_fasta(1), odd_fasta(2)
        p1 = os.path.join(td, "a.fa")
        p2 = os.path.join(td, "b.fa.gz")
        open(p1, "wb").write(t1)
        with gzip.open(p2, "wb") as f:
            f.write(t2)
        r1, n1 = parse(t1)
        r2, n2 = parse(t2)
        recs, names = r1 + r2, n1 + n2
        for k in (31, 63, 15):
            g = cf1g.Cf1g(k=k, num_bins=8)
            st = g.load_fasta([p1, p2], threads=4)
            assert st["n_records"] == len(recs) and st["bases"] == sum(map(len, recs))
            a, off, nm = g.loaded_records()
            assert nm == names
            assert [a[off[i]:off[i + 1]].tobytes() for i in range(len(recs))] == recs
            h = cf1g.Cf1g(k=k, num_bins=8)
            h.load_records(recs)
            h.count_minimizer_histogram(want=False)
            s1, m1 = g.stream_export()
            s2, m2 = h.stream_export()
            np.testing.assert_array_equal(m1, m2, err_msg="placeholder bits")
            np.testing.assert_array_equal(s1, s2, err_msg="2-bit bases")
            check_same(products(g), products(h))
            g.close()
            h.close()


def test_large_fasta_many_threads():
    a, off, names = synth.genome_300m(scale=0.05)
    with tempfile.TemporaryDirectory() as td:
        p = os.path.join(td, "g.fa")
        synth.write_fasta(p, a, off, names, width=61)
        recs = synth.records_list(a, off)
        g = cf1g.Cf1g(k=31, num_bins=64)
        st = g.load_fasta(p, threads=16)
        assert st["bases"] == len(a)
        ga, goff, gn = g.loaded_records()
        assert gn == list(names)
        np.testing.assert_array_equal(goff, np.asarray(off, np.uint64) - off[0])
        assert ga.tobytes() == bytes(a)
        h = cf1g.Cf1g(k=31, num_bins=64)
        h.load_records(recs)
        h.count_minimizer_histogram(want=False)
        for x, y in zip(g.stream_export(), h.stream_export()):
            np.testing.assert_array_equal(x, y)
        check_same(products(g), products(h))
        g.close()
        h.close()


def test_missing_file_and_order():
    g = cf1g.Cf1g(k=31, num_bins=8)
    with pytest.raises(cf1g.Cf1gError) as e:
        g.load_fasta("/nonexistent/x.fa")
    assert e.value.code == -1
    with pytest.raises(cf1g.Cf1gError) as e:
        g.loaded_records()
    assert e.value.code == -2
    g.load_records([b"ACGT" * 20])
    with pytest.raises(cf1g.Cf1gError) as e:
        g.stream_export()  # packed by the device during P1
    assert e.value.code == -2
    g.close()


def test_p5_from_fasta_equals_p5_from_records():
    a, off, names = synth.genome_5m(scale=0.02)
    recs = synth.records_list(a, off)
    with tempfile.TemporaryDirectory() as td:
        p = os.path.join(td, "g.fa")
        synth.write_fasta(p, a, off, names)
        out = []
        for via_fasta in (True, False):
            g = cf1g.Cf1g(k=31, num_bins=16)
            if via_fasta:
                g.load_fasta(p, threads=3)
            else:
                g.load_records(recs)
            g.run_p1_p4()
            pre = os.path.join(td, "o%d" % via_fasta)
            meta = g.p5_write(pre, names=None if via_fasta else list(names), input_desc="g.fa")
            out.append((meta, sorted(open(pre + ".cf_seg").read().splitlines()),
                        sorted(open(pre + ".cf_seq").read().splitlines())))
            g.close()
        assert out[0] == out[1]


def test_cli_build_matches_oracle():
    # the command-line driver: FASTA (plain + gzip) -> P1..P4 on the GPU -> P5
    import json
    import subprocess

    from oracle import RefPipeline

    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    exe = os.path.join(root, "cf1-rs_amd", "bin", "cf1g")
    a, off, names = synth.genome_5m(scale=0.02)
    recs = synth.records_list(a, off)
    with tempfile.TemporaryDirectory() as td:
        p1 = os.path.join(td, "a.fa")
        synth.write_fasta(p1, a, off, names)
        extra = [b"ACGTTGCAAC" * 9, b"ACG", b"NNNNACGTACGTACGTACGTACGTACGTACGTACGTNNNN"]
        p2 = os.path.join(td, "b.fa.gz")
        with gzip.open(p2, "wb") as f:
            for i, s in enumerate(extra):
                f.write(b">x%d desc\n%s\n" % (i, s))
        pre = os.path.join(td, "out")
        r = subprocess.run([exe, "build", "-s", p1, "-s", p2, "-o", pre, "-k", "31", "-t", "4", "--track-short-seqs",
                            "--poly-N-stretch"], capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stderr
        ref = RefPipeline(31, num_bins=128, threads=4)
        ref.load(recs + extra, list(names) + ["x0", "x1", "x2"])
        ref.set_files([0] * len(recs) + [1] * len(extra))
        ref.run_p1_p4()
        rp = os.path.join(td, "ref")
        ref.p5(rp, input_desc=p1 + ", " + p2)
        for ext in (".cf_seg", ".cf_seq"):
            assert sorted(open(pre + ext).read().splitlines()) == sorted(open(rp + ext).read().splitlines()), ext
        # ref_id = file_idx + 1 (src/output.rs:176-177); x1 is shorter than k
        heads = sorted(line.split("\t")[0] for line in open(pre + ".cf_seq").read().splitlines())
        assert heads == sorted(["Reference:1_Sequence:" + n for n in names] + ["Reference:2_Sequence:x0", "Reference:2_Sequence:x2"])
        jg, jr = json.load(open(pre + ".json")), json.load(open(rp + ".json"))
        jg["parameters info"].pop("output prefix")
        jr["parameters info"].pop("output prefix")
        assert jg == jr


def test_two_files_reference_ids():
    # records of the second input file are Reference:2 in cf_seq, through
    # cf1g_load_fasta and through cf1g_load_records + cf1g_set_record_files
    from oracle import RefPipeline

    r = random.Random(4)
    f1 = [rand_seq(r, 400).encode() for _ in range(3)]
    f2 = [rand_seq(r, 300).encode() for _ in range(2)]
    names = ["a0", "a1", "a2", "b0", "b1"]
    with tempfile.TemporaryDirectory() as td:
        paths = [os.path.join(td, "1.fa"), os.path.join(td, "2.fa")]
        for p, recs, nm in ((paths[0], f1, names[:3]), (paths[1], f2, names[3:])):
            with open(p, "wb") as f:
                for n, s in zip(nm, recs):
                    f.write(b">" + n.encode() + b"\n" + s + b"\n")
        ref = RefPipeline(31, num_bins=8, threads=2)
        ref.load(f1 + f2, names)
        ref.set_files([0, 0, 0, 1, 1])
        ref.run_p1_p4()
        rp = os.path.join(td, "ref")
        ref.p5(rp, input_desc="x")
        want = sorted(open(rp + ".cf_seq").read().splitlines())
        assert sum(w.startswith("Reference:2_Sequence:b") for w in want) == 2
        for via in ("fasta", "records"):
            g = cf1g.Cf1g(k=31, num_bins=8)
            if via == "fasta":
                g.load_fasta(paths, threads=2)
            else:
                g.load_records(f1 + f2)
                assert g.record_files().tolist() == [0] * 5
                g.set_record_files([0, 0, 0, 1, 1])
            assert g.record_files().tolist() == [0, 0, 0, 1, 1]
            g.run_p1_p4()
            pre = os.path.join(td, via)
            g.p5_write(pre, names=names, input_desc="x")
            assert sorted(open(pre + ".cf_seq").read().splitlines()) == want, via
            with pytest.raises(cf1g.Cf1gError):
                g.set_record_files([0, 1])
            g.close()
