"""The host codecs, the generator and the oracle under AddressSanitizer and
UndefinedBehaviorSanitizer (tests/asan: a CPU-only build that exercises
DIV-RF encode/decode at several partition, lane and packet settings, every
device layout, packet-group order, lossy rounding, the oracle's conversions
and the Matrix Market fixtures; any sanitizer report aborts it)."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++ with libasan")
def test_host_code_under_asan_ubsan():
    r = subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "tests", "asan")], capture_output=True, text=True,
                       timeout=900)
    print(r.stdout[-2000:], r.stderr[-4000:])
    assert r.returncode == 0
    assert "ok (0 failures)" in r.stdout
