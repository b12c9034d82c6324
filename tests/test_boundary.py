"""The C-ABI library loads here (no GPU) and exports every function that
include/cf1g.h declares; argument and device errors come back as codes."""
import ctypes as C
import os
import re

import pytest

import cf1g

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    txt = open(os.path.join(ROOT, "include", "cf1g.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(cf1g_[a-z0-9_]+)\s*\(", txt)))


def test_header_symbols_exported():
    names = declared()
    assert len(names) >= 20
    L = cf1g.lib()
    for n in names:
        assert hasattr(L, n), n
    assert sorted(cf1g.EXPORTS) == names


def test_invalid_k_and_bins():
    for k in (0, 2, 30, 64, 65):
        with pytest.raises(cf1g.Cf1gError) as e:
            cf1g.Cf1g(k=k)
        assert e.value.code == -1
    for nb in (0, 4097):
        with pytest.raises(cf1g.Cf1gError) as e:
            cf1g.Cf1g(k=31, num_bins=nb)
        assert e.value.code == -1


def test_null_context_is_rejected():
    L = cf1g.lib()
    assert L.cf1g_p1_histogram(None, None) == -1
    assert L.cf1g_run_p1_p4(None, None) == -1
    assert L.cf1g_last_error(None) == b"null context"


def test_no_device_fails_loudly():
    try:
        import torch  # noqa: F401
        have_gpu = torch.cuda.is_available()
    except Exception:
        have_gpu = False
    if have_gpu:
        pytest.skip("a GPU is visible")
    with pytest.raises(cf1g.Cf1gError) as e:
        cf1g.Cf1g(k=31)
    assert e.value.code == -4


def test_minimizer_length_must_fit_records():
    # super k-mers of up to 2k - m bases must fit a routed record (57 bases
    # for k <= 31, 121 for k > 32); refused before any device is touched
    for k, m in ((31, 1), (31, 3), (63, 1), (63, 3), (9, 11)):
        with pytest.raises(cf1g.Cf1gError) as e:
            cf1g.Cf1g(k=k, m=m)
        assert e.value.code == -1
