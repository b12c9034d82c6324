"""The reference's intermediate files written from GPU products, and resume:
cf1_bin_%05d.bin (u16 LE base count + 2-bit bytes per record) holds, bin by
bin, the same super k-mer records as the oracle's routing into the
reference's byte format; cf1_dedup_all.bin holds each bin's distinct keys
sorted, bins in order; cf1g_save / cf1g_restore bring back an MPHF and state
vector whose lookups, states, annotation and P5 output equal the run that
was saved."""
import os
import tempfile

import numpy as np
import pytest

import cf1g
from gpu_cases import edge_records, random_genome
from oracle import RefPipeline
from test_gpu_parity import decode_ref_bin

pytestmark = pytest.mark.gpu


def pipelines(seqs, k, num_bins):
    ref = RefPipeline(k, num_bins=num_bins, threads=4)
    ref.load(seqs)
    ref.p1()
    ref.p2()
    ref.p3()
    ref.p4()
    g = cf1g.Cf1g(k=k, num_bins=num_bins)
    g.load_records(seqs)
    g.run_p1_p4()
    return ref, g


def raw_records(data):
    """Records of a bin file as (count, bytes) pairs, checking the framing."""
    out, p = [], 0
    while p < len(data):
        assert p + 2 <= len(data)
        n = data[p] | (data[p + 1] << 8)
        nb = (n + 3) // 4
        assert p + 2 + nb <= len(data)
        out.append(data[p:p + 2 + nb])
        p += 2 + nb
    return out


@pytest.mark.parametrize("k", [5, 31, 63])
def test_bin_files_match_oracle_routing(k):
    seqs = random_genome(seed=700 + k, n_rec=5, max_len=6000) + edge_records()
    ref, g = pipelines(seqs, k, 16)
    with tempfile.TemporaryDirectory() as td:
        nbytes = g.bins_write(td, threads=4)
        nb = g.info()["num_bins"]
        files = sorted(os.listdir(td))
        assert files == ["cf1_bin_%05d.bin" % b for b in range(nb)]
        total = 0
        for b in range(nb):
            data = open(os.path.join(td, files[b]), "rb").read()
            total += len(data)
            want = ref.bin_bytes(b)
            # same records (the order within a bin follows the schedule)
            assert sorted(raw_records(data)) == sorted(raw_records(want)), f"bin {b}"
            assert sorted(decode_ref_bin(data)) == sorted(decode_ref_bin(want))
        assert total == nbytes
    g.close()


@pytest.mark.parametrize("k", [21, 31, 33, 63])
def test_dedup_file(k):
    seqs = random_genome(seed=800 + k, n_rec=4, max_len=8000)
    ref, g = pipelines(seqs, k, 8)
    with tempfile.TemporaryDirectory() as td:
        p = os.path.join(td, "cf1_dedup_all.bin")
        g.dedup_write(p, threads=3)
        raw = np.fromfile(p, np.uint64)
        lo, hi = g.keys()
        D = len(lo)
        if k > 32:
            assert raw.size == 2 * D
            got = list(zip(raw[1::2].tolist(), raw[0::2].tolist()))  # (hi, lo)
            keyset = set(zip(hi.tolist(), lo.tolist()))
        else:
            assert raw.size == D
            got = [(0, int(x)) for x in raw.tolist()]
            keyset = set((0, int(x)) for x in lo.tolist())
        assert set(got) == keyset and len(set(got)) == D
        # bins in order, each sorted: the bin of each key from the MPHF table
        mw = g.mphf_export()
        sb, ns = int(mw[3]), int(mw[5])
        key_off = mw[16 + 16384:16 + 16384 + ns + 1].astype(np.int64)
        nb = ns >> sb
        for b in range(nb):
            a0, a1 = key_off[b << sb], key_off[(b + 1) << sb]
            run = got[a0:a1]
            assert run == sorted(run), f"bin {b} not sorted"
            # the keys of bin b are those the MPHF gives values [a0, a1)
            vk = list(zip(hi.tolist(), lo.tolist())) if k > 32 else [(0, int(x)) for x in lo.tolist()]
            assert set(run) == set(vk[a0:a1]), f"bin {b} keys"
        # against the oracle's key set
        rlo, rhi = ref.keys()
        want = set(zip(rhi.tolist(), rlo.tolist())) if k > 32 else set((0, int(x)) for x in rlo.tolist())
        assert set(got) == want
    g.close()


@pytest.mark.parametrize("k,via_fasta", [(31, False), (63, False), (31, True)])
def test_save_restore(k, via_fasta):
    seqs = random_genome(seed=900 + k, n_rec=4, max_len=9000) + edge_records()
    g = cf1g.Cf1g(k=k, num_bins=16)
    g.load_records(seqs)
    g.run_p1_p4()
    lo, hi = g.keys()
    want = (g.mphf_export(), g.states(), g.annotate(), g.mphf_lookup(lo, hi))
    with tempfile.TemporaryDirectory() as td:
        path = os.path.join(td, "run.cf1g")
        g.save(path)
        g.close()
        h = cf1g.Cf1g(k=k, num_bins=16)
        if via_fasta:
            fa = os.path.join(td, "r.fa")
            with open(fa, "wb") as f:
                for i, s in enumerate(seqs):
                    f.write(b">r%d\n%s\n" % (i, s))
            h.load_fasta(fa, threads=2)
        else:
            h.load_records(seqs)
        h.restore(path)
        got = (h.mphf_export(), h.states(), h.annotate(), h.mphf_lookup(lo, hi))
        for x, y, what in zip(got, want, ("MPHF", "states", "annotation", "lookups")):
            np.testing.assert_array_equal(x, y, err_msg=what)
        # the CAS classifier works on a restored MPHF and gives the same states
        h.classify_vertices_cas()
        np.testing.assert_array_equal(h.states(), want[1])
        # keys and joined observations are not part of the file
        for f in (h.keys, h.classify_vertices):
            with pytest.raises(cf1g.Cf1gError) as e:
                f()
            assert e.value.code == -2
        # a full run afterwards rebuilds everything
        h.run_p1_p4()
        np.testing.assert_array_equal(h.states(), want[1])
        h.close()
        # other records: refused
        o = cf1g.Cf1g(k=k, num_bins=16)
        o.load_records(seqs[:-1])
        with pytest.raises(cf1g.Cf1gError) as e:
            o.restore(path)
        assert e.value.code == -1
        o.close()


@pytest.mark.parametrize("k", [3, 31, 63])
def test_unitig_spans_partition_runs(k):
    # spans tile every run of valid k-mers, in stream order; each distinct id
    # is claimed exactly once; orientation and id agree with the oracle's P5
    seqs = random_genome(seed=950 + k, n_rec=4, max_len=5000) + edge_records()
    g = cf1g.Cf1g(k=k, num_bins=8)
    g.load_records(seqs)
    g.run_p1_p4()
    s, e, ids, fl = g.unitig_spans()
    ann = g.annotate()
    valid = (ann & 1).astype(bool)
    covered = np.zeros(len(ann), bool)
    assert np.all(s <= e) and np.all(e[:-1] < s[1:])
    for a, b in zip(s.tolist(), e.tolist()):
        covered[a:b + 1] = True
    np.testing.assert_array_equal(covered, valid)
    claimed = ids[(fl & 2) != 0]
    assert len(set(claimed.tolist())) == len(claimed) == len(set(ids.tolist()))
    g.close()


def test_unitig_spans_follow_new_records():
    # the spans a count query marks are kept for the fill that follows it;
    # new records, a new P4 or another index must mark them again
    def ref_lines(seqs, k):
        ref = RefPipeline(k, num_bins=8, threads=4)
        ref.load(seqs)
        for ph in (ref.p1, ref.p2, ref.p3, ref.p4):
            ph()
        with tempfile.TemporaryDirectory() as td:
            ref.p5(os.path.join(td, "r"), input_desc="in.fa")
            return sorted(open(os.path.join(td, "r.cf_seg")).read().splitlines())

    def gpu_lines(g):
        with tempfile.TemporaryDirectory() as td:
            g.p5_write(os.path.join(td, "g"), input_desc="in.fa")
            return sorted(open(os.path.join(td, "g.cf_seg")).read().splitlines())

    a = random_genome(seed=961, n_rec=3, max_len=4000)
    b = random_genome(seed=962, n_rec=5, max_len=2500) + edge_records()
    g = cf1g.Cf1g(k=31, num_bins=8)
    g.load_records(a)
    g.run_p1_p4()
    n_a = len(g.unitig_spans()[0])
    assert gpu_lines(g) == ref_lines(a, 31)
    g.load_records(b)
    g.run_p1_p4()
    assert gpu_lines(g) == ref_lines(b, 31)
    g.classify_vertices_cas()
    assert gpu_lines(g) == ref_lines(b, 31)
    g.load_records(a)
    g.run_p1_p4()
    assert len(g.unitig_spans()[0]) == n_a
    g.close()


def _digest(words):
    """Save-body checksum: sum mod 2^64 of mix64(word ^ index * golden)."""
    M = (1 << 64) - 1
    z = (words ^ (np.arange(len(words), dtype=np.uint64) * np.uint64(0x9E3779B97F4A7C15)))
    with np.errstate(over="ignore"):
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return int(z.sum(dtype=np.uint64)) & M


def test_restore_rejects_damaged_files():
    seqs = random_genome(seed=990, n_rec=2, max_len=3000)
    g = cf1g.Cf1g(k=31, num_bins=4)
    g.load_records(seqs)
    g.run_p1_p4()
    with tempfile.TemporaryDirectory() as td:
        path = os.path.join(td, "run.cf1g")
        g.save(path)
        data = open(path, "rb").read()
        w = np.frombuffer(data, np.uint64).copy()
        mw = int(w[6])
        ns = int(w[16 + 5])
        tab = 65536 // 4
        ko = 16 + 16 + tab          # first key offset word of the MPHF body
        bad = []
        bad.append(data[: len(data) - 8])                       # truncated
        hdr = bytearray(data)
        hdr[6 * 8:7 * 8] = (1 << 50).to_bytes(8, "little")      # absurd size in the header
        bad.append(bytes(hdr))
        bad.append(b"not a save file" + data[15:])              # wrong magic
        for word in (16 + 20, ko + 1, ko + ns + 2, 16 + mw - 3, 16 + mw + 1):  # any body word flipped
            x = w.copy()
            x[word] ^= np.uint64(1 << 33)
            bad.append(x.tobytes())

        def resealed(x):  # a damaged body with a matching checksum: the structure checks
            x[9] = np.uint64(_digest(x[16:]))
            return x.tobytes()

        x = w.copy()
        x[16 + 16] = np.uint64(0xFFFF)                          # bucket->bin entries beyond the bins
        bad.append(resealed(x))
        x = w.copy()
        x[ko + 1], x[ko + 2] = x[ko + 2], x[ko + 1]             # key offsets not monotone
        bad.append(resealed(x))
        x = w.copy()
        x[ko + 2 * ns + 1] += np.uint64(1)                      # last block offset != block count
        bad.append(resealed(x))
        x = w.copy()
        x[16 + 7] = np.uint64(1 << 62)                          # block count that would wrap the size arithmetic
        bad.append(resealed(x))
        for i, b in enumerate(bad):
            p = os.path.join(td, "bad%d" % i)
            open(p, "wb").write(b)
            h = cf1g.Cf1g(k=31, num_bins=4)
            h.load_records(seqs)
            with pytest.raises(cf1g.Cf1gError) as e:
                h.restore(p)
            assert e.value.code == -1, (i, e.value)
            h.close()
        # records of the same count and length, other bases: refused (stream digest)
        other = [bytes(b"ACGT"[(b"ACGT".index(c) + 1) % 4] if c in b"ACGT" else c for c in s) for s in seqs]
        h = cf1g.Cf1g(k=31, num_bins=4)
        h.load_records(other)
        with pytest.raises(cf1g.Cf1gError) as e:
            h.restore(path)
        assert e.value.code == -1 and "digest" in str(e.value)
        h.close()
        # the untouched file still restores
        h = cf1g.Cf1g(k=31, num_bins=4)
        h.load_records(seqs)
        h.restore(path)
        np.testing.assert_array_equal(h.states(), g.states())
        h.close()
    g.close()
