"""ctypes view of liboracle.so -- TEST INFRASTRUCTURE ONLY.

Imported by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg,
as the checker; the product (spmv-dare_amd/) never imports it.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_path = os.path.join(HERE, "liboracle.so")
if not os.path.exists(_path):
    raise ImportError("%s not built: run `make -C oracle`" % _path)
_lib = ctypes.CDLL(_path)
_vp, _l = ctypes.c_void_p, ctypes.c_long
_sig = {
    "oracle_csr_naive": [_vp, _vp, _vp, _vp, _vp, _l],
    "oracle_csr_dd": [_vp, _vp, _vp, _vp, _vp, _vp, _l],
    "oracle_csr_naive_omp": [_vp, _vp, _vp, _vp, _vp, _l, ctypes.c_int],
    "oracle_partition_iterations": [_l, _l, _l, _l, _vp, _vp],
    "oracle_partition_prefix_sums": [_l, _l, _vp, _l, ctypes.c_int64, _vp, _vp],
    "oracle_sellcs_order": [_vp, _l, _l, _vp],
    "oracle_sellcs_layout": [_vp, _l, _l, _vp, ctypes.c_int, _vp, _vp],
    "oracle_sellcs_fill": [_vp, _vp, _vp, _l, _l, _vp, _vp, _vp, _vp, _vp],
    "oracle_sellcs_spmv": [_l, _l, _vp, _vp, _vp, _vp, _vp, _vp, _vp],
    "oracle_sell_layout": [_vp, _l, _l, _vp],
    "oracle_sell_fill": [_vp, _vp, _vp, _l, _l, _l, _vp, _vp, _vp],
    "oracle_sell_spmv": [_l, _l, _vp, _vp, _vp, _vp, _vp],
    "oracle_sellsorted_windows": [_vp, _l, _l, _l, _vp],
    "oracle_sellsorted_order": [_vp, _l, _l, _vp, _vp],
    "oracle_sellsorted_layout": [_vp, _l, _l, _vp, _vp],
    "oracle_sellsorted_fill": [_vp, _vp, _vp, _l, _l, _l, _vp, _vp, _vp, _vp, _vp],
    "oracle_sellsorted_spmv": [_l, _l, _vp, _vp, _vp, _vp, _vp, _vp],
    "oracle_div_reduce_precision": [ctypes.c_double, ctypes.c_double],
    "oracle_fnv1a": [_vp, _l],
}
_res = {"oracle_csr_naive_omp": ctypes.c_int, "oracle_sellcs_layout": _l, "oracle_sellcs_fill": _l,
        "oracle_sell_layout": _l, "oracle_sellsorted_layout": _l, "oracle_fnv1a": ctypes.c_uint64,
        "oracle_div_reduce_precision": ctypes.c_double}
for _n, _a in _sig.items():
    getattr(_lib, _n).argtypes = _a
    getattr(_lib, _n).restype = _res.get(_n)


def _c(a, dt):
    a = np.ascontiguousarray(a, dtype=dt)
    return a, a.ctypes.data


def csr_naive(rp, ci, va, x):
    rp, p0 = _c(rp, np.int32)
    ci, p1 = _c(ci, np.int32)
    va, p2 = _c(va, np.float64)
    x, p3 = _c(x, np.float64)
    m = rp.size - 1
    y = np.empty(m, dtype=np.float64)
    _lib.oracle_csr_naive(p0, p1, p2, p3, y.ctypes.data, m)
    return y


def csr_naive_omp(rp, ci, va, x, threads=0, y=None):
    rp, p0 = _c(rp, np.int32)
    ci, p1 = _c(ci, np.int32)
    va, p2 = _c(va, np.float64)
    x, p3 = _c(x, np.float64)
    m = rp.size - 1
    if y is None:
        y = np.empty(m, dtype=np.float64)
    t = _lib.oracle_csr_naive_omp(p0, p1, p2, p3, y.ctypes.data, m, threads)
    return y, t


def csr_dd(rp, ci, va, x):
    """(hi, lo) double-double row sums."""
    rp, p0 = _c(rp, np.int32)
    ci, p1 = _c(ci, np.int32)
    va, p2 = _c(va, np.float64)
    x, p3 = _c(x, np.float64)
    m = rp.size - 1
    hi = np.empty(m, dtype=np.float64)
    lo = np.empty(m, dtype=np.float64)
    _lib.oracle_csr_dd(p0, p1, p2, p3, hi.ctypes.data, lo.ctypes.data, m)
    return hi, lo


def partition_iterations(W, w, start, end):
    s, e = ctypes.c_long(), ctypes.c_long()
    _lib.oracle_partition_iterations(W, w, start, end, ctypes.byref(s), ctypes.byref(e))
    return s.value, e.value


def partition_prefix_sums(W, w, sums, total):
    sums, p = _c(sums, np.int64)
    s, e = ctypes.c_long(), ctypes.c_long()
    _lib.oracle_partition_prefix_sums(W, w, p, sums.size, int(total), ctypes.byref(s), ctypes.byref(e))
    return s.value, e.value


def sellcs_convert(rp, ci, va, C, sigma, width_rule=0):
    """Reference-semantics SELL-C-sigma arrays (int64 indices, as INDEX64)."""
    rp, p0 = _c(rp, np.int32)
    ci, p1 = _c(ci, np.int32)
    va, p2 = _c(va, np.float64)
    m = rp.size - 1
    ns = (m + C - 1) // C
    order = np.empty(m, dtype=np.int64)
    _lib.oracle_sellcs_order(p0, m, sigma, order.ctypes.data)
    widths = np.empty(ns, dtype=np.int64)
    sptr = np.empty(ns + 1, dtype=np.int64)
    total = _lib.oracle_sellcs_layout(p0, m, C, order.ctypes.data, width_rule, widths.ctypes.data,
                                      sptr.ctypes.data)
    cols = np.empty(max(total, 1), dtype=np.int64)[:total]
    vals = np.empty(max(total, 1), dtype=np.float64)[:total]
    trunc = _lib.oracle_sellcs_fill(p0, p1, p2, m, C, order.ctypes.data, widths.ctypes.data,
                                    sptr.ctypes.data, cols.ctypes.data, vals.ctypes.data)
    return dict(order=order, widths=widths, slice_ptr=sptr, cols=cols, vals=vals, truncated=trunc)


def sellcs_spmv(conv, C, x):
    m = conv["order"].size
    x, px = _c(x, np.float64)
    y = np.zeros(m, dtype=np.float64)
    _lib.oracle_sellcs_spmv(m, C, conv["order"].ctypes.data, conv["widths"].ctypes.data,
                            conv["slice_ptr"].ctypes.data, conv["cols"].ctypes.data,
                            conv["vals"].ctypes.data, px, y.ctypes.data)
    return y


def sell_convert(rp, ci, va, VL, T=1):
    rp, p0 = _c(rp, np.int32)
    ci, p1 = _c(ci, np.int32)
    va, p2 = _c(va, np.float64)
    m = rp.size - 1
    nc = (m + VL - 1) // VL
    cptr = np.empty(nc + 1, dtype=np.int64)
    total = _lib.oracle_sell_layout(p0, m, VL, cptr.ctypes.data)
    ja = np.empty(max(total, 1), dtype=np.int64)[:total]
    a = np.empty(max(total, 1), dtype=np.float64)[:total]
    _lib.oracle_sell_fill(p0, p1, p2, m, VL, T, cptr.ctypes.data, ja.ctypes.data, a.ctypes.data)
    return dict(cluster_ptr=cptr, ja=ja, a=a)


def sell_spmv(conv, VL, x, m):
    x, px = _c(x, np.float64)
    y = np.zeros(m, dtype=np.float64)
    _lib.oracle_sell_spmv(m, VL, conv["cluster_ptr"].ctypes.data, conv["ja"].ctypes.data,
                          conv["a"].ctypes.data, px, y.ctypes.data)
    return y


def sellsorted_windows(rp, VL, T):
    """Row starts of the reference's T thread windows (T + 1 entries)."""
    rp, p0 = _c(rp, np.int32)
    out = np.empty(T + 1, dtype=np.int64)
    _lib.oracle_sellsorted_windows(p0, rp.size - 1, VL, T, out.ctypes.data)
    return out


def sellsorted_convert(rp, ci, va, VL, windows):
    """SELL_SORTED over the given row windows (row starts, last = m)."""
    rp, p0 = _c(rp, np.int32)
    ci, p1 = _c(ci, np.int32)
    va, p2 = _c(va, np.float64)
    win, pw = _c(windows, np.int64)
    m = rp.size - 1
    perm = np.empty(max(m, 1), dtype=np.int32)[:m]
    _lib.oracle_sellsorted_order(p0, m, win.size - 1, pw, perm.ctypes.data)
    nc = (m + VL - 1) // VL
    cptr = np.empty(nc + 1, dtype=np.int64)
    total = _lib.oracle_sellsorted_layout(p0, m, VL, perm.ctypes.data, cptr.ctypes.data)
    ja = np.empty(max(total, 1), dtype=np.int64)[:total]
    a = np.empty(max(total, 1), dtype=np.float64)[:total]
    _lib.oracle_sellsorted_fill(p0, p1, p2, m, VL, win.size - 1, pw, perm.ctypes.data, cptr.ctypes.data,
                                ja.ctypes.data, a.ctypes.data)
    order = np.empty(m, dtype=np.int32)
    order[perm] = np.arange(m, dtype=np.int32)
    return dict(perm=perm, order=order, cluster_ptr=cptr, ja=ja, a=a)


def sellsorted_spmv(conv, VL, x):
    x, px = _c(x, np.float64)
    m = conv["perm"].size
    y = np.zeros(m, dtype=np.float64)
    _lib.oracle_sellsorted_spmv(m, VL, conv["cluster_ptr"].ctypes.data, conv["ja"].ctypes.data,
                                conv["a"].ctypes.data, conv["perm"].ctypes.data, px, y.ctypes.data)
    return y


def div_reduce_precision(values, tolerance):
    return np.array([_lib.oracle_div_reduce_precision(float(v), float(tolerance)) for v in np.ravel(values)])


def fnv1a(a):
    a = np.ascontiguousarray(a)
    return "%016x" % _lib.oracle_fnv1a(a.ctypes.data, a.nbytes)


def normwise_rel_err(y, y_ref):
    """||y - y_ref||_2 / ||y_ref||_2 (0 when both are zero)."""
    d = np.linalg.norm(np.asarray(y, dtype=np.float64) - np.asarray(y_ref, dtype=np.float64))
    r = np.linalg.norm(y_ref)
    return float(d / r) if r > 0 else float(d)


def dd_rel_err(y, hi, lo):
    """Normwise error of y against the double-double y = hi + lo."""
    d = (np.asarray(y) - hi) - lo
    r = np.linalg.norm(hi)
    return float(np.linalg.norm(d) / r) if r > 0 else float(np.linalg.norm(d))


_lib.oracle_div_decode.argtypes = [_vp, _l, _l, _vp, _vp, _vp, _l]
_lib.oracle_div_decode.restype = _l


def div_decode(stream, L, nnz):
    """(rows, cols, vals) of a DIV-RF packet stream, packet by packet."""
    b = np.ascontiguousarray(stream, dtype=np.uint8)
    rows = np.empty(max(nnz, 1), dtype=np.int32)
    cols = np.empty(max(nnz, 1), dtype=np.int32)
    vals = np.empty(max(nnz, 1), dtype=np.float64)
    k = _lib.oracle_div_decode(b.ctypes.data, b.size, L, rows.ctypes.data, cols.ctypes.data,
                               vals.ctypes.data, nnz)
    if k < 0:
        raise ValueError("malformed DIV stream")
    return rows[:k], cols[:k], vals[:k]


def div_packet_sizes(stream, L):
    """Byte length of each packet, read from the headers."""
    b = np.ascontiguousarray(stream, dtype=np.uint8)
    out = []
    at = 0
    while at < b.size:
        h = b[at:at + 28 + 4 * L].view(np.uint32) if (28 + 4 * L) % 4 == 0 else None
        nv, nu, nrfs = int(h[0]), int(h[1]), int(h[6])
        rb, cbits = int(b[at + 20]), int(b[at + 21])
        cb = (rb + cbits + 7) >> 3
        lanes = int(h[7:7 + L].astype(np.int64).sum())
        size = 28 + 4 * L + 8 * nrfs + nv * cb + nu + lanes
        out.append(size)
        at += size
    return out


def div_mask_padding(stream, L):
    """Copy of a stream with each header's two padding bytes (offsets 22-23)
    zeroed: the reference leaves them uninitialised."""
    b = np.array(stream, dtype=np.uint8, copy=True)
    at = 0
    for s in div_packet_sizes(b, L):
        b[at + 22:at + 24] = 0
        at += s
    return b
