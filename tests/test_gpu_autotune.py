"""The autotuner (cf1g_autotune): every candidate is measured, the context
keeps the fastest, and the products of the chosen configuration equal the
oracle's for the same bin count and segment target."""
import numpy as np
import pytest

import cf1g
from gpu_cases import random_genome
from oracle import RefPipeline

pytestmark = pytest.mark.gpu


def test_autotune_choice_and_parity():
    seqs = random_genome(seed=61, n_rec=6, max_len=20000)
    g = cf1g.Cf1g(k=31, num_bins=16)
    g.load_records(seqs)
    res = g.autotune(reps=1)
    steps = [e for e in res["table"] if e["kind"] == "step"]
    cas = [e for e in res["table"] if e["kind"] == "p4_cas"]
    assert len(steps) == 9 and len(cas) == 9 and res["n_entries"] == 18
    assert res["num_bins"] in (128, 256, 512) and res["p3_target_seg"] in (1024, 1536, 2048)
    assert min(e["ms"] for e in steps) == pytest.approx(res["step_ms"], abs=1e-3)
    assert min(e["ms"] for e in cas) == pytest.approx(res["p4_cas_ms"], abs=1e-3)
    ref = RefPipeline(31, num_bins=res["num_bins"], threads=4, target_seg=res["p3_target_seg"])
    ref.load(seqs)
    d_ref = ref.run_p1_p4()
    # the context holds the chosen configuration's products ...
    assert g.info()["n_distinct"] == d_ref
    np.testing.assert_array_equal(g.mphf_export(), ref.mphf_export())
    np.testing.assert_array_equal(g.states(), ref.states())
    # ... and later runs use it
    assert g.run_p1_p4() == d_ref
    np.testing.assert_array_equal(g.mphf_export(), ref.mphf_export())
    st = g.classify_vertices_cas()
    assert (st["chunk"], st["wgs_per_cu"]) == (res["p4_chunk"], res["p4_wgs_per_cu"])
    np.testing.assert_array_equal(g.states(), ref.states())
    g.close()


def test_autotune_needs_records():
    g = cf1g.Cf1g(k=31, num_bins=16)
    with pytest.raises(cf1g.Cf1gError) as e:
        g.autotune()
    assert e.value.code == -2
    g.close()
