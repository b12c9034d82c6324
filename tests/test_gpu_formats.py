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


def test_restore_rejects_damaged_files():
    seqs = random_genome(seed=990, n_rec=2, max_len=3000)
    g = cf1g.Cf1g(k=31, num_bins=4)
    g.load_records(seqs)
    g.run_p1_p4()
    with tempfile.TemporaryDirectory() as td:
        path = os.path.join(td, "run.cf1g")
        g.save(path)
        data = open(path, "rb").read()
        bad = []
        bad.append(data[: len(data) - 8])                       # truncated
        hdr = bytearray(data)
        hdr[6 * 8:7 * 8] = (1 << 50).to_bytes(8, "little")      # absurd size in the header
        bad.append(bytes(hdr))
        bad.append(b"not a save file" + data[15:])              # wrong magic
        for i, b in enumerate(bad):
            p = os.path.join(td, "bad%d" % i)
            open(p, "wb").write(b)
            h = cf1g.Cf1g(k=31, num_bins=4)
            h.load_records(seqs)
            with pytest.raises(cf1g.Cf1gError) as e:
                h.restore(p)
            assert e.value.code == -1
            h.close()
    g.close()
