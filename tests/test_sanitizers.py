"""Host C++ under AddressSanitizer (+UBSan) and ThreadSanitizer, on the CPU.

tests/sanitize/san_driver.cpp links the product's host code (FASTA ingest
with its reader threads, host packing, the host thread pool the streaming
load uses, the bin / dedup file writers, the P5 line formatter with its
worker threads) and the CPU oracle (its parallel P1..P5) into one program,
built twice by tests/sanitize/Makefile. Each build runs every part with one
thread and with eight and compares the two; the test passes when both
builds exit 0 and neither sanitizer reported anything. The default oracle
input is 0.6 Mbp; CF1G_SAN_BASES=4600000 adds a record across the 4 Mbp
chunk edge (about three minutes per build).
"""
import os
import shutil
import subprocess

import pytest

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "sanitize")
BASES = int(os.environ.get("CF1G_SAN_BASES", "600000"))


def _have(flag):
    if not shutil.which("g++"):
        return False
    r = subprocess.run(["g++", flag, "-x", "c++", "-", "-o", os.devnull], input="int main(){return 0;}",
                       capture_output=True, text=True)
    return r.returncode == 0


@pytest.fixture(scope="module")
def built():
    if not (_have("-fsanitize=address") and _have("-fsanitize=thread")):
        pytest.skip("g++ without the address / thread sanitizer runtimes")
    r = subprocess.run(["make", "-j8"], cwd=HERE, capture_output=True, text=True, timeout=900)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    return os.path.join(HERE, "_build")


@pytest.mark.parametrize("kind", ["asan", "tsan"])
def test_host_code_under_sanitizer(built, kind, tmp_path):
    env = dict(os.environ)
    env["ASAN_OPTIONS"] = "detect_leaks=1:abort_on_error=0"
    env["UBSAN_OPTIONS"] = "print_stacktrace=1"
    env["TSAN_OPTIONS"] = "halt_on_error=1"
    r = subprocess.run([os.path.join(built, f"san_{kind}"), str(tmp_path), "8", str(BASES)], capture_output=True, text=True,
                       env=env, timeout=900)
    out = r.stdout + r.stderr
    assert r.returncode == 0, out[-6000:]
    for bad in ("ERROR: AddressSanitizer", "ERROR: LeakSanitizer", "runtime error:", "WARNING: ThreadSanitizer"):
        assert bad not in out, out[-6000:]
    for part in ("ingest ok", "formats ok", "p5 ok", "oracle ok"):
        assert part in r.stdout
