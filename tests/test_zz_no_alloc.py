"""The allocation-free structure statistics (spmvhip_csr_stats_device_noalloc)
against the allocating entry point (spmvhip_csr_stats_device) through the C
ABI: bit-identical statistics on two calls in a row, the second on a smaller
matrix (the shared scratch keeps the first call's partials), and the row
limit of the scratch rejected by the host before any launch."""
import numpy as np
import pytest

from spmvdare import gen, hip

pytestmark = pytest.mark.gpu

LINES = ["300000 300000 12 6 gamma random 0.2 8 2 0.3 5",   # 74 workgroups of partials
         "5000 7000 9 3 normal diagonal 0.1 2 1 0.1 6"]     # 2 workgroups


@pytest.fixture(scope="module", autouse=True)
def device():
    hip.init(0)


def test_noalloc_stats_match_allocating_entry_point():
    d_out = hip.DeviceArray(6)
    for line in LINES:
        print("case no-alloc stats", line, flush=True)
        a = gen.generate(gen.GenParams.parse(line), stats=False)
        rp = hip.DeviceArray.from_host(a.row_ptr)
        ci = hip.DeviceArray.from_host(a.col_ind)
        ref = hip.csr_stats_device(rp, ci, a.m, a.n)
        d_out.fill_bytes(0xff)
        hip.csr_stats_device_noalloc(rp, ci, a.m, a.n, d_out)
        hip.synchronize()
        got = d_out.download()
        want = np.array([ref[k] for k in ("avg_nnz_per_row", "std_nnz_per_row", "max_nnz_per_row",
                                          "avg_bw_scaled", "avg_num_neighbours", "nonempty_rows")])
        np.testing.assert_array_equal(got, want)
        lens = np.diff(a.row_ptr)
        assert got[2] == lens.max() and got[5] == np.count_nonzero(lens)


def test_noalloc_stats_rejects_rows_beyond_scratch():
    rp = hip.DeviceArray.from_host(np.zeros(2, np.int32))
    ci = hip.DeviceArray.from_host(np.zeros(1, np.int32))
    d_out = hip.DeviceArray(6)
    with pytest.raises(hip.SpmvHipError) as e:
        # refused on the host from m alone: no kernel is launched
        hip.csr_stats_device_noalloc(rp, ci, 67108864 + 1, 10, d_out)
    assert e.value.code == -1 and "exceed" in str(e.value)
