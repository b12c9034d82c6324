"""Shared test setup: import paths for the product bindings (cf1-rs_amd/cf1g)
and the CPU oracle (oracle/), the `gpu` marker, and small input builders."""
import os
import random
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(ROOT, "cf1-rs_amd"), os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) device")


def rand_seq(rng, n, pn=0.0, plow=0.0, alphabet="ACGT"):
    out = []
    for _ in range(n):
        if pn and rng.random() < pn:
            out.append("N")
            continue
        c = rng.choice(alphabet)
        if plow and rng.random() < plow:
            c = c.lower()
        out.append(c)
    return "".join(out)


def rc(s):
    comp = {"A": "T", "C": "G", "G": "C", "T": "A", "a": "T", "c": "G", "g": "C", "t": "A"}
    return "".join(comp.get(c, "N") for c in reversed(s))


@pytest.fixture
def rng():
    return random.Random(12345)
