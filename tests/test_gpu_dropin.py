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


@pytest.mark.parametrize("fmt", ["csr_naive", "csr_vec", "sell", "sell_c_s", "div_rf", "sell_sorted"])
def test_own_driver(fmt):
    exe = os.path.join(ROOT, "spmv-dare_amd", "bin", "spmv_hip_%s_d.exe" % fmt)
    r = subprocess.run([exe] + LINE, env=ENV, capture_output=True, text=True, timeout=120)
    print(r.stdout[-800:], r.stderr[-400:])
    assert r.returncode == 0, r.stderr
    assert "GFLOPS" in r.stdout
    assert "normwise rel err" in r.stdout
    csv = r.stderr.strip().splitlines()[-1].split(",")
    assert csv[18].startswith("HIP_")


@pytest.mark.parametrize("fmt", ["csr_vec", "sell_c_s", "div_rf", "sell_sorted"])
def test_reference_driver(fmt):
    exe = os.path.join(ROOT, "oracle", "_ref", "spmv_hip_%s_d.exe" % fmt)
    if not os.path.exists(exe):
        pytest.skip("reference driver not built here (needs the reference sources at build time)")
    r = subprocess.run([exe] + LINE, env=ENV, capture_output=True, text=True, timeout=120)
    print(r.stdout[-800:], r.stderr[-400:])
    assert r.returncode == 0, r.stderr
    assert "GFLOPS" in r.stdout
    assert "HIP_" in r.stderr  # format_name in the reference's CSV row (bench_spmv.cpp:496)


# Matrix Market files (tests/golden/mtx, pinned to the reference's own reader
# by tests/test_mtx.py): the validation mode, where the driver checks y
# against a compensated _Float128 y_gold (bench_spmv.cpp:109-236)
MTX_DIR = os.path.join(ROOT, "tests", "golden", "mtx")
MTX_FILES = ["general_real", "symmetric_real", "skew_real", "symmetric_pattern", "general_integer",
             "general_complex", "hermitian_complex", "no_banner", "empty_rows"]
ENV_MTX = dict(ENV, USE_ARTIFICIAL_MATRICES="0")
# The reference's driver reads y for check_accuracy right after its timing
# loop (bench_spmv.cpp:434) with no call into the format before it, so its
# validation runs use GPU_KERNEL=0: the drop-in copies y back after every call
# (with GPU_KERNEL=1, y reaches the host at statistics_print_data and at
# destruction, which this driver calls only after the check).
ENV_MTX_REF = dict(ENV_MTX, GPU_KERNEL="0")


def accuracy_line(out):
    line = [l for l in out.splitlines() if l.startswith("errors spmv:")]
    assert line, out[-800:]
    vals = dict(kv.split("=") for kv in line[-1][len("errors spmv: "):].split(", "))
    return {k: float(v) for k, v in vals.items()}


@pytest.mark.parametrize("fmt", ["csr_naive", "csr_vec", "sell", "sell_c_s", "div_rf", "sell_sorted"])
@pytest.mark.parametrize("name", MTX_FILES)
def test_own_driver_matrix_market(fmt, name):
    exe = os.path.join(ROOT, "spmv-dare_amd", "bin", "spmv_hip_%s_d.exe" % fmt)
    path = os.path.join(MTX_DIR, name + ".mtx")
    print("case mtx", fmt, name, flush=True)
    r = subprocess.run([exe, path], env=ENV_MTX, capture_output=True, text=True, timeout=120)
    print(r.stdout[-600:], r.stderr[-300:])
    assert r.returncode == 0, r.stdout[-800:] + r.stderr[-400:]
    assert "Test failed" not in r.stdout
    acc = accuracy_line(r.stdout)
    assert acc["max_ae"] <= 1e-12 and acc["mape"] <= 1e-8


@pytest.mark.parametrize("fmt", ["csr_vec", "sell_c_s", "div_rf", "sell_sorted"])
@pytest.mark.parametrize("name", MTX_FILES)
def test_reference_driver_matrix_market(fmt, name):
    """The reference's unmodified driver reads the file with its own import
    path and checks the HIP format's y with its own check_accuracy."""
    exe = os.path.join(ROOT, "oracle", "_ref", "spmv_hip_%s_d.exe" % fmt)
    if not os.path.exists(exe):
        pytest.skip("reference driver not built here (needs the reference sources at build time)")
    path = os.path.join(MTX_DIR, name + ".mtx")
    print("case mtx reference driver", fmt, name, flush=True)
    r = subprocess.run([exe, path], env=ENV_MTX_REF, capture_output=True, text=True, timeout=120)
    print(r.stdout[-600:], r.stderr[-300:])
    assert r.returncode == 0, r.stdout[-800:] + r.stderr[-400:]
    assert "Test failed" not in r.stdout
    acc = accuracy_line(r.stdout)
    assert acc["max_ae"] <= 1e-12


@pytest.mark.parametrize("fmt", ["csr_naive", "csr_vec", "sell_c_s", "div_rf", "sell_sorted"])
@pytest.mark.parametrize("name", ["general_real", "symmetric_real", "general_complex", "empty_rows"])
def test_reference_cpu_formats_matrix_market(fmt, name):
    """The CPU baseline's executables (the reference's formats and driver;
    sell_c_s with the portable lane-loop kernel) pass the reference's own
    accuracy check."""
    exe = os.path.join(ROOT, "oracle", "_ref", "cpu_%s_d.exe" % fmt)
    if not os.path.exists(exe):
        pytest.skip("reference CPU executables not built here")
    env = dict(ENV_MTX, GPU_KERNEL="0", CSRCV_NUM_PACKET_VALS="16384", DIV_VC_TOLERANCE="0",
               DIV_PACKET_REORDERING="0", DIV_STEALING_ENABLED="1")  # read by the reference's DIV (config.sh)
    r = subprocess.run([exe, os.path.join(MTX_DIR, name + ".mtx")], env=env, capture_output=True, text=True,
                       timeout=120)
    print(r.stdout[-600:], r.stderr[-300:])
    assert r.returncode == 0
    assert "Test failed" not in r.stdout
    assert accuracy_line(r.stdout)["max_ae"] <= 1e-12
