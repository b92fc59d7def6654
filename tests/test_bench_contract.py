"""bench.py's host-side contract (CPU only): the HBM bytes per launch that
its roofline object reports come from the committed rocprofv3 summary, and
that summary still holds a row, with calibrated bytes, for every profiled
variant; the algorithmic bytes of the 8M-row matrix; the one-GPU refusal."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

import bench  # noqa: E402  (the repository root is on the path from conftest; bench.py runs nothing at import)


def test_profiled_traffic_rows_present():
    for (workload, fmt), params in bench.PROFILED.items():
        t = bench.profiled_traffic(workload, fmt, params)
        assert isinstance(t, int) and t > 0, (workload, fmt)
        # calibrated bytes per launch: at least the format's compulsory
        # reads (x and y alone are 134 MB on the 8M-row matrix)
        assert t > 134_000_000
    # div_rf's packets are compressed: fewer bytes than csr_vec's stream
    dv = bench.profiled_traffic("spmv8m", "div_rf", bench.PROFILED[("spmv8m", "div_rf")])
    cv = bench.profiled_traffic("spmv8m", "csr_vec", bench.PROFILED[("spmv8m", "csr_vec")])
    assert dv < cv


def test_profiled_traffic_other_variant_is_none():
    assert bench.profiled_traffic("spmv8m", "div_rf", {"div_packet_vals": 512}) is None
    assert bench.profiled_traffic("spmv100k", "csr_vec", {"block_nnz": 256}) is None


def test_alg_bytes_8m():
    # 12 nnz + 4 (m + 1) + 8 n + 8 m for the 8,388,608-row matrix of 268,448,832 nonzeros
    assert bench.alg_bytes(8388608, 8388608, 268448832) == 3389158148


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
