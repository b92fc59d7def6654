"""DIV-RF on the GPU: the device decoder's (row, col, value) sequence is
bit-identical to the oracle's decode of the same packet stream (whose encoder
is byte-identical to the reference's, tests/test_div.py), and y is within
1e-13 normwise of csr_naive."""
import numpy as np
import pytest

from spmvdare import gen, hip
import pyoracle
from test_gpu_formats import matrix, run, check_dd, TOL

pytestmark = pytest.mark.gpu

NAMES = ["one_row_empty", "all_empty", "block_edges", "huge_rows", "ragged_4097", "test_txt", "cfg1",
         "batch_03", "dense_rows", "tall", "degree_clamp"]
CODECS = [(1, 64, 16384), (4, 4, 512), (2, 32, 4096), (3, 64, 300)]


@pytest.fixture(scope="module", autouse=True)
def device():
    hip.init(0)


@pytest.mark.parametrize("name", NAMES)
@pytest.mark.parametrize("T,L,P", CODECS)
@pytest.mark.parametrize("unroll", [4, 8, 16])
def test_div_device_decode_and_spmv(name, T, L, P, unroll):
    """unroll picks the decoder instance: 4, 8 or 16 values per lane and step."""
    rp, ci, va, n = matrix(name)
    m = rp.size - 1
    print("case", name, T, L, P, unroll, flush=True)
    f = hip.csr_to_format(rp, ci, va, m, n, ci.size, kind="div_rf", div_threads=T, div_lanes=L,
                          div_packet_vals=P, unroll=unroll)
    if ci.size:
        b, off, grp = hip.div_encode(rp, ci, va, n, T=T, L=L, P=P)
        r, c, v = pyoracle.div_decode(b, L, ci.size)
        np.testing.assert_array_equal(f.export(6, np.int32), r)
        np.testing.assert_array_equal(f.export(7, np.int32), c)
        np.testing.assert_array_equal(f.export(8, np.float64), v)
    x = gen.make_x(n, 91)
    y = run(f, x, m)
    assert pyoracle.normwise_rel_err(y, pyoracle.csr_naive(rp, ci, va, x)) <= TOL
    assert (y[np.diff(rp) == 0] == 0).all()
    check_dd(y, rp, ci, va, x, "div_rf T%d L%d P%d/%s" % (T, L, P, name))


def test_div_lanes_above_wave_rejected():
    rp, ci, va, n = matrix("test_txt")
    with pytest.raises(hip.SpmvHipError) as e:
        hip.csr_to_format(rp, ci, va, rp.size - 1, n, ci.size, kind="div_rf", div_lanes=128)
    assert e.value.code == -1
