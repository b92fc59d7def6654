"""The HIP Matrix_Format objects as drop-ins: driven by this repository's
driver executables (one per format, as the reference links them) and by the
reference's own, unmodified driver (bench.cpp + bench_spmv.cpp) compiled
against the reference's spmv_kernel.h and linked with the HIP object
(oracle/Makefile target `refdrop`; skipped where it was not built)."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LINE = "20000 20000 12 9 gamma diagonal 0.05 5 2 0.3 3".split()
# the variables config.sh exports and the reference driver reads with atoi(getenv(...))
ENV = dict(os.environ, USE_ARTIFICIAL_MATRICES="1", USE_PROCESSES="0", GPU_KERNEL="1", USE_RCM_REORDERING="0",
           COOLDOWN="0", NUM_PROCESSES="1", PROGG="spmv_hip", SPMV_MIN_LOOPS="16", SPMV_MIN_RUNTIME="0.05",
           OMP_NUM_THREADS="4")


@pytest.mark.parametrize("fmt", ["csr_naive", "csr_vec", "sell", "sell_c_s", "div_rf"])
def test_own_driver(fmt):
    exe = os.path.join(ROOT, "spmv-dare_amd", "bin", "spmv_hip_%s_d.exe" % fmt)
    r = subprocess.run([exe] + LINE, env=ENV, capture_output=True, text=True, timeout=120)
    print(r.stdout[-800:], r.stderr[-400:])
    assert r.returncode == 0, r.stderr
    assert "GFLOPS" in r.stdout
    assert "normwise rel err" in r.stdout
    csv = r.stderr.strip().splitlines()[-1].split(",")
    assert csv[18].startswith("HIP_")


@pytest.mark.parametrize("fmt", ["csr_vec", "sell_c_s", "div_rf"])
def test_reference_driver(fmt):
    exe = os.path.join(ROOT, "oracle", "_ref", "spmv_hip_%s_d.exe" % fmt)
    if not os.path.exists(exe):
        pytest.skip("reference driver not built here (needs the reference sources at build time)")
    r = subprocess.run([exe] + LINE, env=ENV, capture_output=True, text=True, timeout=120)
    print(r.stdout[-800:], r.stderr[-400:])
    assert r.returncode == 0, r.stderr
    assert "GFLOPS" in r.stdout
    assert "HIP_" in r.stderr  # format_name in the reference's CSV row (bench_spmv.cpp:496)
