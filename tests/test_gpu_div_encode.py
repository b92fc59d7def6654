"""DIV-RF packets encoded on the device (csrc/hip/div_enc.hip) against the
host encoder, whose packets are byte-identical to the reference's own build
(tests/test_div.py): the device-layout bytes and the packet offsets of both
encoders must be identical -- reference layout (1) and interleaved lane
streams (3), lossless and lossy, over partition counts, lane counts and packet
sizes -- and the two handles must give the same y."""
import numpy as np
import pytest

from spmvdare import gen, hip
from test_gpu_formats import matrix

pytestmark = pytest.mark.gpu

NAMES = ["block_edges", "huge_rows", "ragged_4097", "single_column", "test_txt", "cfg1", "batch_03",
         "dense_rows", "tall", "degree_clamp"]
WIDE = {"wide_cols_4m": "20000 4000000 16 4 normal random 0.5 1.5 1 0.2 3",
        "wide_diag_2m": "50000 2000000 20 8 gamma diagonal 0.2 10 3 0.4 9"}
CODECS = [(1, 64, 1024), (4, 4, 512), (2, 32, 2048), (3, 64, 300), (1, 64, 1)]


@pytest.fixture(scope="module", autouse=True)
def device():
    hip.init(0)


def load(name):
    if name in WIDE:
        a = gen.generate(gen.GenParams.parse(WIDE[name]), stats=False)
        return a.row_ptr, a.col_ind, a.values, a.n
    return matrix(name)


def compare(rp, ci, va, n, label, **kw):
    m = rp.size - 1
    print("case", label, kw, flush=True)
    fh = hip.csr_to_format(rp, ci, va, m, n, ci.size, kind="div_rf", div_encoder=1, **kw)
    fd = hip.csr_to_format(rp, ci, va, m, n, ci.size, kind="div_rf", div_encoder=2, **kw)
    off_h, off_d = fh.export(11, np.int64), fd.export(11, np.int64)
    np.testing.assert_array_equal(off_d, off_h)
    b_h, b_d = fh.export(10, np.uint8), fd.export(10, np.uint8)
    assert b_h.size == b_d.size
    diff = np.flatnonzero(b_h != b_d)
    if diff.size:
        p = int(np.searchsorted(off_h, diff[0], side="right") - 1)
        raise AssertionError("%s: %d bytes differ, first at %d (packet %d at %d)" % (
            label, diff.size, diff[0], p, off_h[p]))
    ih, idv = fh.info(), fd.info()
    for k in ("work_items", "padded_nnz"):
        if k in ih:
            assert ih[k] == idv[k], k
    x = gen.make_x(n, 17)
    dx = hip.DeviceArray.from_host(x)
    dy1, dy2 = hip.DeviceArray(max(m, 1)), hip.DeviceArray(max(m, 1))
    fh.spmv_device(dx, dy1)
    fd.spmv_device(dx, dy2)
    hip.synchronize()
    np.testing.assert_array_equal(dy1.download(), dy2.download())
    fh.destroy()
    fd.destroy()
    return b_d.size


@pytest.mark.parametrize("name", NAMES + sorted(WIDE))
@pytest.mark.parametrize("T,L,P", CODECS)
@pytest.mark.parametrize("layout", [1, 3])
def test_device_encoder_bytes(name, T, L, P, layout):
    rp, ci, va, n = load(name)
    if ci.size == 0:
        pytest.skip("no stored entry: no packets")
    if P == 1 and ci.size > 200000:
        pytest.skip("one-value packets: kept to the smaller matrices")
    compare(rp, ci, va, n, name, div_threads=T, div_lanes=L, div_packet_vals=P, div_layout=layout)


@pytest.mark.parametrize("name", ["cfg1", "dense_rows", "wide_diag_2m"])
@pytest.mark.parametrize("T,L,P", [(2, 64, 1024), (3, 32, 300)])
def test_device_encoder_repeated_values(name, T, L, P):
    """Values rounded to halves: replication-factor groups above 1 in every
    packet (runs of equal values, the counting sort by replication factor,
    unit-wise interleaving of the coordinates)."""
    rp, ci, va, n = load(name)
    compare(rp, ci, np.round(va * 2) / 2, n, name + "_halves", div_threads=T, div_lanes=L, div_packet_vals=P)


@pytest.mark.parametrize("name", ["cfg1", "batch_03", "wide_cols_4m"])
@pytest.mark.parametrize("tol", [0.3, 1e-3, 1e-9])
def test_device_encoder_lossy(name, tol):
    """DIV_VC_TOLERANCE: the values rounded on the device as on the host."""
    rp, ci, va, n = load(name)
    compare(rp, ci, va, n, "%s_tol%g" % (name, tol), div_threads=2, div_packet_vals=1024, div_tolerance=tol)


def test_device_encoder_automatic_and_limits():
    """The automatic choice encodes packets of at most 2048 values on the
    device; larger packets, packet reordering and the x-slot layouts stay on
    the host, and asking for the device encoder there is an error."""
    rp, ci, va, n = matrix("cfg1")
    m = rp.size - 1
    for kw in (dict(div_packet_vals=4096), dict(div_packet_vals=1024, div_reorder=1),
               dict(div_packet_vals=1024, div_layout=2)):
        with pytest.raises(hip.SpmvHipError):
            hip.csr_to_format(rp, ci, va, m, n, ci.size, kind="div_rf", div_encoder=2, **kw)
        hip.csr_to_format(rp, ci, va, m, n, ci.size, kind="div_rf", **kw).destroy()
    with pytest.raises(hip.SpmvHipError):
        hip.csr_to_format(rp, ci, va, m, n, ci.size, kind="div_rf", div_encoder=3)


def test_device_encoder_from_device_arrays():
    """csr_to_format_device: the CSR stays on the device (only the row
    pointers come over for the packet plan) and gives the host path's bytes."""
    rp, ci, va, n = matrix("batch_03")
    m = rp.size - 1
    fh = hip.csr_to_format(rp, ci, va, m, n, ci.size, kind="div_rf", div_packet_vals=1024, div_encoder=1)
    d_rp, d_ci, d_va = (hip.DeviceArray.from_host(a) for a in (rp, ci, va))
    fd = hip.csr_to_format_device(d_rp, d_ci, d_va, m, n, ci.size, kind="div_rf", div_packet_vals=1024)
    np.testing.assert_array_equal(fd.export(10, np.uint8), fh.export(10, np.uint8))
    np.testing.assert_array_equal(fd.export(11, np.int64), fh.export(11, np.int64))
