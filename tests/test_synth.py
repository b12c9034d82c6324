"""Synthetic inputs have the documented shapes (scaled-down versions)."""
import numpy as np

from cf1g import synth


def test_genome_5m_shape():
    a, off, names = synth.genome_5m(scale=0.02)
    assert len(names) == 1 and int(off[-1]) == 100000
    assert set(np.unique(a).tolist()) <= set(b"ACGT")


def test_transcriptome_shape():
    a, off, names = synth.transcriptome(scale=0.01, k=31)
    lens = np.diff(off.astype(np.int64))
    assert len(lens) == 1000
    assert 0.005 < (lens < 31).mean() < 0.05
    assert lens.max() <= 100000
    assert 700 < np.median(lens[lens >= 31]) < 2000


def test_genome_300m_shape():
    a, off, names = synth.genome_300m(scale=0.02)
    lens = np.diff(off.astype(np.int64))
    assert len(lens) == 24 and lens.sum() == 6_000_000
    assert abs(lens.max() - 800_000) < 20_000
    acgt = a[a != ord("N")]
    gc = np.isin(acgt, [ord("C"), ord("G")]).mean()
    assert 0.37 < gc < 0.45
    assert (a == ord("N")).sum() > 0 or lens.max() <= 5_000_000


def test_deterministic():
    a1, o1, _ = synth.transcriptome(scale=0.002)
    a2, o2, _ = synth.transcriptome(scale=0.002)
    assert np.array_equal(a1, a2) and np.array_equal(o1, o2)
