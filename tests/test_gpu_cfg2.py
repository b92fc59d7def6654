"""Parity at full size on the 8,388,608-row generated matrix (avg 32 nnz/row,
gamma std 48, skew 20, diagonal placement, bw 0.05, 4 neighbours, cross-row
similarity 0.5, seed 7; 268,448,832 nnz), with the exact format parameters
bench.py times on it.

Bars: every format's y within 1e-13 normwise of the oracle's csr_naive y
(sequential left-to-right row sums), its error against the double-double
oracle printed; the DIV-RF device decode of the whole packet stream
(268M (row, col, value) triples) bit-identical to the oracle's decode of the
same stream, whose encoder is byte-identical to the reference's."""
import json
import os

import numpy as np
import pytest

from spmvdare import gen, hip
import pyoracle

pytestmark = [pytest.mark.gpu, pytest.mark.timeout(900)]
HERE = os.path.dirname(os.path.abspath(__file__))
LINE = "8388608 8388608 32 48 gamma diagonal 0.05 20 4 0.5 7"
TOL = 1e-13

# the variants bench.py reports on this matrix, plus the reference layouts
VARIANTS = [
    ("div_rf", dict(div_packet_vals=1024, div_threads=64, unroll=8, xcd_remap=1)),
    ("csr_vec", dict(block_nnz=256)),
    ("csr_vec", dict(block_nnz=256, xcd_remap=1)),
    ("csr_vec", dict(block_nnz=256, xcd_remap=1, narrow_loads=1)),
    ("sell_c_s", dict(sigma=256, lanes_per_row=4, unroll=4)),
    ("sell_c_s", dict(sigma=256)),
    ("sell_c_s", dict(sigma=4096, lanes_per_row=4, unroll=4)),
    ("sell_c_s", dict(sigma=1)),
    ("sell", dict(lanes_per_row=4, unroll=4)),
]


@pytest.fixture(scope="module")
def cfg2():
    hip.init(0)
    p = gen.GenParams.parse(LINE)
    a = gen.generate(p, stats=False)
    golden = json.load(open(os.path.join(HERE, "golden", "generator.json")))["cfg2"]
    assert (a.m, a.n, a.nnz) == (golden["m"], golden["n"], golden["nnz"])
    x = gen.make_x(a.n, p.seed + 1)
    y_ref, _ = pyoracle.csr_naive_omp(a.row_ptr, a.col_ind, a.values, x)
    hi, lo = pyoracle.csr_dd(a.row_ptr, a.col_ind, a.values, x)
    print("cfg2: m=%d nnz=%d; csr_naive vs double-double %.3e" % (
        a.m, a.nnz, pyoracle.dd_rel_err(y_ref, hi, lo)), flush=True)
    dx = hip.DeviceArray.from_host(x)
    dy = hip.DeviceArray(a.m)
    return a, x, y_ref, hi, lo, dx, dy


@pytest.mark.parametrize("kind,kw", VARIANTS, ids=["%s-%s" % (k, "-".join("%s%s" % i for i in sorted(kw.items())))
                                                    for k, kw in VARIANTS])
def test_cfg2_y(cfg2, kind, kw):
    a, x, y_ref, hi, lo, dx, dy = cfg2
    print("case cfg2", kind, kw, flush=True)
    f = hip.csr_to_format(a.row_ptr, a.col_ind, a.values, a.m, a.n, a.nnz, kind=kind, **kw)
    dy.fill_bytes(0xff)  # NaN in every row the kernel leaves unwritten
    f.spmv_device(dx, dy)
    f.spmv_device(dx, dy)  # a second launch over the first one's y: no state carried between calls
    hip.synchronize()
    y = dy.download()
    err = pyoracle.normwise_rel_err(y, y_ref)
    dd = pyoracle.dd_rel_err(y, hi, lo)
    print("cfg2 %s %s: normwise vs csr_naive %.3e, vs double-double %.3e" % (kind, kw, err, dd), flush=True)
    assert np.isfinite(y).all()
    assert err <= TOL
    assert dd <= TOL
    f.destroy()


def test_cfg2_div_decode_bit_exact(cfg2):
    """The whole device-side decode of the bench's DIV-RF stream (P = 1024,
    64 partitions, 64 lanes) against the oracle's decode of the host stream."""
    a = cfg2[0]
    T, L, P = 64, 64, 1024
    print("case cfg2 div decode", T, L, P, flush=True)
    f = hip.csr_to_format(a.row_ptr, a.col_ind, a.values, a.m, a.n, a.nnz, kind="div_rf", div_threads=T,
                          div_lanes=L, div_packet_vals=P, unroll=8, xcd_remap=1)
    b, off, grp = hip.div_encode(a.row_ptr, a.col_ind, a.values, a.n, T=T, L=L, P=P)
    r, c, v = pyoracle.div_decode(b, L, a.nnz)
    del b
    np.testing.assert_array_equal(f.export(6, np.int32), r)
    del r
    np.testing.assert_array_equal(f.export(7, np.int32), c)
    del c
    np.testing.assert_array_equal(f.export(8, np.float64), v)
    f.destroy()
