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


# matrices with wide column spans: their packets have 4-byte coordinates
# (row bits rounded to 8 plus up to 24 column bits; the decoder's aligned
# 64-lane path), 5-byte coordinates, or a mix of widths
WIDE = {"wide_cols_4m": "20000 4000000 16 4 normal random 0.5 1.5 1 0.2 3",
        "wide_cols_30m": "30000 30000000 24 12 gamma random 0.9 30 2 0.3 5",
        "wide_diag_2m": "50000 2000000 20 8 gamma diagonal 0.2 10 3 0.4 9"}


@pytest.mark.parametrize("name", NAMES)
@pytest.mark.parametrize("T,L,P", CODECS)
@pytest.mark.parametrize("unroll", [4, 8, 16])
def test_div_device_decode_and_spmv(name, T, L, P, unroll):
    """unroll picks the decoder instance: 4, 8 or 16 values per lane and step."""
    check_div(*matrix(name), T, L, P, unroll, name)


@pytest.mark.parametrize("name", sorted(WIDE))
@pytest.mark.parametrize("T,L,P", [(1, 64, 16384), (3, 64, 300), (2, 64, 1024), (2, 32, 1024)])
@pytest.mark.parametrize("unroll", [4, 8, 16])
def test_div_wide_columns(name, T, L, P, unroll):
    a = gen.generate(gen.GenParams.parse(WIDE[name]), stats=False)
    check_div(a.row_ptr, a.col_ind, a.values, a.n, T, L, P, unroll, name)


@pytest.mark.parametrize("name", ["wide_cols_4m", "wide_diag_2m"])
@pytest.mark.parametrize("copies", [1, 2, 64])
def test_div_accumulator_copies(name, copies):
    """The row accumulators' copy count (automatic elsewhere) changes only
    which LDS address a lane adds into."""
    a = gen.generate(gen.GenParams.parse(WIDE[name]), stats=False)
    check_div(a.row_ptr, a.col_ind, a.values, a.n, 2, 64, 1024, 8, name, div_acc_copies=copies)


@pytest.mark.parametrize("name", ["cfg1", "batch_03", "dense_rows", "wide_cols_4m", "wide_diag_2m"])
@pytest.mark.parametrize("T,L,P", [(2, 64, 1024), (3, 32, 300)])
@pytest.mark.parametrize("unroll", [4, 16])
def test_div_repeated_values(name, T, L, P, unroll):
    """Values rounded to multiples of 1/2 repeat within packets, so every
    packet holds replication-factor groups above 1 (the decoders' general
    path, which the generated matrices' distinct values hardly reach), with
    1- to 4-byte coordinates."""
    if name in WIDE:
        a = gen.generate(gen.GenParams.parse(WIDE[name]), stats=False)
        rp, ci, va, n = a.row_ptr, a.col_ind, a.values, a.n
    else:
        rp, ci, va, n = matrix(name)
    check_div(rp, ci, np.round(va * 2) / 2, n, T, L, P, unroll, name + "_halves")


@pytest.mark.parametrize("name", ["wide_diag_2m", "ragged_4097", "dense_rows"])
@pytest.mark.parametrize("remap", [1, 3, 100000])
def test_div_xcd_chunks(name, remap):
    """Packet groups dealt to the XCDs in chunks (64, 3, or one chunk larger
    than the whole grid) on a padded grid: every group is decoded once."""
    if name in WIDE:
        a = gen.generate(gen.GenParams.parse(WIDE[name]), stats=False)
        rp, ci, va, n = a.row_ptr, a.col_ind, a.values, a.n
    else:
        rp, ci, va, n = matrix(name)
    check_div(rp, ci, va, n, 2, 64, 1024, 8, name, xcd_remap=remap)


def test_div_accumulator_copies_rejected():
    rp, ci, va, n = matrix("test_txt")
    for bad in (3, 128, -1):
        with pytest.raises(hip.SpmvHipError):
            hip.csr_to_format(rp, ci, va, rp.size - 1, n, ci.size, kind="div_rf", div_acc_copies=bad)


def check_div(rp, ci, va, n, T, L, P, unroll, name, **kw):
    m = rp.size - 1
    print("case", name, T, L, P, unroll, kw, flush=True)
    f = hip.csr_to_format(rp, ci, va, m, n, ci.size, kind="div_rf", div_threads=T, div_lanes=L,
                          div_packet_vals=P, unroll=unroll, **kw)
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


@pytest.mark.parametrize("name", ["test_txt", "cfg1", "ragged_4097", "dense_rows", "huge_rows", "wide_diag_2m"])
@pytest.mark.parametrize("layout", [1, 2, 3, 4])
@pytest.mark.parametrize("T,L,P,unroll", [(2, 64, 1024, 8), (3, 32, 300, 4), (1, 64, 512, 16)])
def test_div_device_layouts(name, layout, T, L, P, unroll):
    """Every device layout (1 reference bytes, 2 x slots, 3 interleaved lane
    streams, 4 both): device decode bit-identical to the oracle's, y within
    1e-13 of csr_naive."""
    if name in WIDE:
        a = gen.generate(gen.GenParams.parse(WIDE[name]), stats=False)
        rp, ci, va, n = a.row_ptr, a.col_ind, a.values, a.n
    else:
        rp, ci, va, n = matrix(name)
    check_div(rp, ci, va, n, T, L, P, unroll, name, div_layout=layout)


@pytest.mark.parametrize("name", ["cfg1", "dense_rows"])
@pytest.mark.parametrize("layout", [2, 3, 4])
def test_div_device_layouts_repeated_values(name, layout):
    """Replication-factor groups above 1 (values rounded to halves) in every
    device layout: the decoders' general path."""
    rp, ci, va, n = matrix(name)
    check_div(rp, ci, np.round(va * 2) / 2, n, 2, 64, 1024, 8, name + "_halves", div_layout=layout)


def test_div_layout_rejected():
    rp, ci, va, n = matrix("test_txt")
    for bad in (-1, 5):
        with pytest.raises(hip.SpmvHipError):
            hip.csr_to_format(rp, ci, va, rp.size - 1, n, ci.size, kind="div_rf", div_layout=bad)


@pytest.mark.parametrize("name", ["test_txt", "cfg1", "batch_03", "ragged_4097", "huge_rows", "dense_rows",
                                  "one_row_empty", "wide_diag_2m"])
@pytest.mark.parametrize("T,P,layout", [(1, 16384, 0), (1, 16384, 1), (4, 512, 1), (4, 512, 2), (4, 512, 3),
                                        (3, 300, 0), (3, 300, 4)])
def test_div_packet_reordering(name, T, P, layout):
    """DIV_PACKET_REORDERING=1: the groups run in the encoder's median-column
    order (checked against the reference's reordered streams on the CPU);
    every group owns its rows, so y is bit-identical to stream order."""
    rp, ci, va, n = wide_or_matrix(name)
    m = rp.size - 1
    print("case reorder", name, layout, T, P, flush=True)
    x = gen.make_x(n, 91)
    kw = dict(div_threads=T, div_packet_vals=P, div_layout=layout, unroll=8)
    f0 = hip.csr_to_format(rp, ci, va, m, n, ci.size, kind="div_rf", **kw)
    f1 = hip.csr_to_format(rp, ci, va, m, n, ci.size, kind="div_rf", div_reorder=1, **kw)
    y0, y1 = run(f0, x, m), run(f1, x, m)
    np.testing.assert_array_equal(y1, y0)
    check_dd(y1, rp, ci, va, x, "div_rf/reorder/T%d/P%d/layout%d/%s" % (T, P, layout, name))


def wide_or_matrix(name):
    if name in WIDE:
        a = gen.generate(gen.GenParams.parse(WIDE[name]), stats=False)
        return a.row_ptr, a.col_ind, a.values, a.n
    return matrix(name)


@pytest.mark.parametrize("name", ["test_txt", "cfg1", "batch_03", "dense_rows", "huge_rows", "one_row_empty"])
@pytest.mark.parametrize("tol", [1e-3, 1e-6, 1e-9])
@pytest.mark.parametrize("layout", [0, 1])
def test_div_lossy_values(name, tol, layout):
    """DIV_VC_TOLERANCE: the device y is the product of the rounded values
    (within 1e-13 normwise of the oracle's csr_naive over the oracle's
    rounding, which is pinned to the reference's lossy streams), and within
    the tolerance of the exact product."""
    rp, ci, va, n = wide_or_matrix(name)
    m = rp.size - 1
    print("case lossy", name, tol, layout, flush=True)
    x = gen.make_x(n, 89)
    f = hip.csr_to_format(rp, ci, va, m, n, ci.size, kind="div_rf", div_tolerance=tol, div_layout=layout,
                          div_packet_vals=4096)
    y = run(f, x, m)
    vr = pyoracle.div_reduce_precision(va, tol)
    y_ref = pyoracle.csr_naive(rp, ci, vr, x)
    err = pyoracle.normwise_rel_err(y, y_ref)
    print("lossy %s tol=%g: normwise vs rounded-value csr_naive %.3e" % (name, tol, err), flush=True)
    assert err <= TOL
    # |y - A x| <= tol |A| |x| entrywise (every value moved by at most tol of itself)
    y_exact = pyoracle.csr_naive(rp, ci, va, x)
    y_abs = pyoracle.csr_naive(rp, ci, np.abs(va), np.abs(x))
    assert np.linalg.norm(y - y_exact) <= tol * np.linalg.norm(y_abs) * 1.001 + 1e-300


@pytest.mark.parametrize("name", ["cfg1", "dense_rows", "huge_rows", "batch_03"])
@pytest.mark.parametrize("layout,copies", [(1, 1), (1, 0), (3, 0), (3, 4), (2, 2)])
def test_div_repeated_launches_bit_identical(name, layout, copies):
    """The products reach their row accumulator in LDS through atomic adds
    (ds_add_f64); lanes of one instruction that hit the same accumulator are
    applied by the LDS in a fixed lane order, instructions in program order,
    and the accumulator copies are summed in a fixed order, so repeated
    launches give the same bits. Checked over 8 launches per case."""
    rp, ci, va, n = wide_or_matrix(name)
    m = rp.size - 1
    print("case repeat", name, layout, copies, flush=True)
    x = gen.make_x(n, 87)
    f = hip.csr_to_format(rp, ci, va, m, n, ci.size, kind="div_rf", div_layout=layout, div_acc_copies=copies,
                          div_packet_vals=512 if layout == 2 else 4096)
    dx = hip.DeviceArray.from_host(x)
    dy = hip.DeviceArray(max(m, 1))
    first = None
    for _ in range(8):
        f.spmv_device(dx, dy)
        hip.synchronize()
        y = dy.download()[:m]
        if first is None:
            first = y
        else:
            assert np.array_equal(y.view(np.uint64), first.view(np.uint64))
    check_dd(first, rp, ci, va, x, "div_rf/repeat/layout%d/copies%d/%s" % (layout, copies, name))
