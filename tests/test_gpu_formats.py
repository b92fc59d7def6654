"""Parity of every HIP format with the CPU oracle, through the C ABI.

Bars (stated per check):
  csr_naive        y bit-identical to the oracle's sequential multiply-then-add
  sell, sell_c_s,  converted arrays bit-identical to the oracle's conversion
  sell_sorted      (which is pinned to the reference's own conversion by the
                   golden fixtures) and y bit-identical to the oracle's
                   lane-sequential fused multiply-add
  csr_vec          normwise relative error <= 1e-13 against csr_naive
every y is also checked against the double-double oracle (error printed).
"""
import functools
import json
import os

import numpy as np
import pytest

from spmvdare import gen, hip
import pyoracle

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
GEN = json.load(open(os.path.join(HERE, "golden", "generator.json")))
TOL = 1e-13


@pytest.fixture(scope="module", autouse=True)
def device():
    hip.init(0)
    print("device:", hip.device_name(), flush=True)


def csr_from_lengths(lens, n, seed):
    rng = np.random.default_rng(seed)
    lens = np.minimum(np.asarray(lens, dtype=np.int64), n)
    rp = np.zeros(lens.size + 1, dtype=np.int32)
    rp[1:] = np.cumsum(lens)
    cols = [np.sort(rng.choice(n, size=int(k), replace=False)) if k else np.zeros(0, np.int64) for k in lens]
    ci = (np.concatenate(cols) if cols else np.zeros(0)).astype(np.int32)
    va = rng.uniform(-1, 1, size=ci.size)
    return rp, ci, va


def generated(name):
    a = gen.generate(gen.GenParams.parse(GEN[name]["line"]), stats=False)
    return a.row_ptr, a.col_ind, a.values, a.n


def edge_cases():
    out = {}
    out["one_row_empty"] = (np.array([0, 0], np.int32), np.zeros(0, np.int32), np.zeros(0), 5)
    rp, ci, va = csr_from_lengths([0] * 200, 7, 1)
    out["all_empty"] = (rp, ci, va, 7)
    rp, ci, va = csr_from_lengths([3, 0, 1, 64, 65, 0, 511, 512, 513, 1023, 1024, 1025, 2, 0] * 13, 3000, 2)
    out["block_edges"] = (rp, ci, va, 3000)
    lens = np.full(3000, 4)
    lens[[5, 700, 2999]] = [40000, 9000, 20000]
    rp, ci, va = csr_from_lengths(lens, 50000, 3)
    out["huge_rows"] = (rp, ci, va, 50000)
    rp, ci, va = csr_from_lengths(np.ones(1000, np.int64), 1, 4)
    out["single_column"] = (rp, ci, va, 1)
    rng = np.random.default_rng(5)
    rp, ci, va = csr_from_lengths(rng.integers(0, 30, 4097), 10000, 5)
    out["ragged_4097"] = (rp, ci, va, 10000)
    return out


CASES = {**edge_cases()}
GEN_CASES = ["test_txt", "cfg1", "batch_03", "batch_07", "stream_01", "run_sh_2", "tall", "wide",
             "dense_rows", "degree_clamp"]


@functools.lru_cache(maxsize=None)
def matrix(name):
    if name in CASES:
        return CASES[name]
    return generated(name)


def run(fmt, x, m):
    """y through the host-pointer contract (y pre-filled with 1.0 like
    src/bench_spmv.cpp:561-564) and through the device path; both must agree."""
    y_host = np.full(m, 1.0)
    fmt.spmv(x, y_host)
    dx = hip.DeviceArray.from_host(x)
    dy = hip.DeviceArray(max(m, 1))
    dy.upload(np.full(max(m, 1), 7.0))
    fmt.spmv_device(dx, dy)
    hip.synchronize()
    y_dev = dy.download()[:m]
    np.testing.assert_array_equal(y_host, y_dev)
    return y_host


def check_dd(y, rp, ci, va, x, label):
    hi, lo = pyoracle.csr_dd(rp, ci, va, x)
    err = pyoracle.dd_rel_err(y, hi, lo)
    print("%s: normwise error vs double-double oracle = %.3e" % (label, err), flush=True)
    assert err <= TOL
    return err


ALL = sorted(CASES) + GEN_CASES


@pytest.mark.parametrize("name", ALL)
def test_csr_naive_bit_exact(name):
    rp, ci, va, n = matrix(name)
    m = rp.size - 1
    print("case", name, m, n, ci.size, flush=True)
    x = gen.make_x(n, 99)
    f = hip.csr_to_format(rp, ci, va, m, n, ci.size, kind="csr_naive")
    y = run(f, x, m)
    np.testing.assert_array_equal(y, pyoracle.csr_naive(rp, ci, va, x))
    check_dd(y, rp, ci, va, x, "csr_naive/" + name)


CSR_VEC_VARIANTS = [dict(), dict(lanes_per_row=1), dict(block_nnz=256, split_nnz=2048),
                    dict(block_nnz=1024), dict(block_nnz=1024, split_nnz=65536), dict(xcd_remap=1),
                    dict(pipeline=1), dict(pipeline=1, block_nnz=512, split_nnz=2048),
                    dict(pipeline=1, block_nnz=128, lanes_per_row=4)]


@pytest.mark.parametrize("name", ALL)
@pytest.mark.parametrize("vi", range(len(CSR_VEC_VARIANTS)))
def test_csr_vec(name, vi):
    rp, ci, va, n = matrix(name)
    m = rp.size - 1
    params = CSR_VEC_VARIANTS[vi]
    print("case", name, params, flush=True)
    x = gen.make_x(n, 98)
    f = hip.csr_to_format(rp, ci, va, m, n, ci.size, kind="csr_vec", **params)
    y = run(f, x, m)
    y0 = pyoracle.csr_naive(rp, ci, va, x)
    assert pyoracle.normwise_rel_err(y, y0) <= TOL
    assert (y[np.diff(rp) == 0] == 0).all()
    check_dd(y, rp, ci, va, x, "csr_vec%s/%s" % (params, name))
    if vi == 0 and m > 0:
        blocks = f.export(5, np.int32)
        assert blocks[0] == 0 and blocks[-1] == m and (np.diff(blocks) > 0).all()
        assert (np.diff(blocks) <= 64).all()


@pytest.mark.parametrize("name", ALL)
@pytest.mark.parametrize("sigma", [64, 256, 4096, 16384])
def test_sell_c_sigma_bit_exact(name, sigma):
    rp, ci, va, n = matrix(name)
    m = rp.size - 1
    print("case", name, sigma, flush=True)
    x = gen.make_x(n, 97)
    f = hip.csr_to_format(rp, ci, va, m, n, ci.size, kind="sell_c_s", sigma=sigma)
    ref = pyoracle.sellcs_convert(rp, ci, va, 64, sigma, width_rule=0)
    if m > 0:
        np.testing.assert_array_equal(f.export(0, np.int32), ref["order"])
    np.testing.assert_array_equal(f.export(1, np.int32), ref["widths"])
    np.testing.assert_array_equal(f.export(2, np.int64), ref["slice_ptr"])
    np.testing.assert_array_equal(f.export(3, np.int32), ref["cols"])
    np.testing.assert_array_equal(f.export(4, np.float64), ref["vals"])
    y = run(f, x, m)
    np.testing.assert_array_equal(y, pyoracle.sellcs_spmv(ref, 64, x))
    check_dd(y, rp, ci, va, x, "sell_c_s/s%d/%s" % (sigma, name))


@pytest.mark.parametrize("name", ALL)
@pytest.mark.parametrize("sigma,y_mode", [(64, 0), (256, 0), (4096, 0), (16384, 0), (4096, 3), (256, 1)])
def test_sell_sorted_bit_exact(name, sigma, y_mode):
    """SELL_SORTED (sell_sorted.cpp) with windows of sigma rows: the
    permutation, slices (including the padding columns) and y bit-identical
    to the oracle over the same windows (the oracle is pinned to the
    reference's own build over its thread windows)."""
    rp, ci, va, n = matrix(name)
    m = rp.size - 1
    print("case sell_sorted", name, sigma, y_mode, flush=True)
    x = gen.make_x(n, 95)
    f = hip.csr_to_format(rp, ci, va, m, n, ci.size, kind="sell_sorted", sigma=sigma, y_mode=y_mode)
    assert f.info()["format_name"] == "HIP_SELL_SORTED"
    ref = pyoracle.sellsorted_convert(rp, ci, va, 64, list(range(0, m, sigma)) + [m])
    if m > 0:
        np.testing.assert_array_equal(f.export(0, np.int32), ref["order"])
    np.testing.assert_array_equal(f.export(2, np.int64), ref["cluster_ptr"])
    np.testing.assert_array_equal(f.export(1, np.int32), np.diff(ref["cluster_ptr"]) // 64)
    np.testing.assert_array_equal(f.export(3, np.int32), ref["ja"])
    np.testing.assert_array_equal(f.export(4, np.float64), ref["a"])
    y = run(f, x, m)
    np.testing.assert_array_equal(y, pyoracle.sellsorted_spmv(ref, 64, x))
    check_dd(y, rp, ci, va, x, "sell_sorted/s%d/%s" % (sigma, name))


def test_sell_sorted_window_must_be_whole_slices():
    rp, ci, va, n = matrix("test_txt")
    with pytest.raises(hip.SpmvHipError, match="multiple of 64"):
        hip.csr_to_format(rp, ci, va, rp.size - 1, n, ci.size, kind="sell_sorted", sigma=100)


@pytest.mark.parametrize("name", ["test_txt", "cfg1", "ragged_4097", "huge_rows"])
@pytest.mark.parametrize("sigma", [1, 100])
def test_sell_c_sigma_unaligned_window(name, sigma):
    """sigma % 64 != 0 is ill-defined in the reference (the slice width is
    the first sorted row, sellcs_format.c:152, and sigma = 1 aborts it);
    here the width is the slice's longest row -- oracle width rule 1."""
    rp, ci, va, n = matrix(name)
    m = rp.size - 1
    x = gen.make_x(n, 96)
    f = hip.csr_to_format(rp, ci, va, m, n, ci.size, kind="sell_c_s", sigma=sigma)
    ref = pyoracle.sellcs_convert(rp, ci, va, 64, sigma, width_rule=1)
    np.testing.assert_array_equal(f.export(0, np.int32), ref["order"])
    np.testing.assert_array_equal(f.export(2, np.int64), ref["slice_ptr"])
    np.testing.assert_array_equal(f.export(3, np.int32), ref["cols"])
    y = run(f, x, m)
    np.testing.assert_array_equal(y, pyoracle.sellcs_spmv(ref, 64, x))


@pytest.mark.parametrize("name", ["test_txt", "cfg1", "ragged_4097", "huge_rows", "block_edges", "one_row_empty",
                                  "run_sh_2", "dense_rows"])
@pytest.mark.parametrize("kind,T", [("sell_c_s", 2), ("sell_c_s", 8), ("sell_c_s", 64), ("sell", 4)])
def test_sell_lanes_per_row(name, kind, T):
    """T lanes per row: same permutation and slices, entries regrouped so that
    a wave step reads T consecutive entries of 64 / T rows; T = 1 is the
    reference layout (bit-exact tests above), T > 1 changes the summation
    order, so y is checked normwise."""
    rp, ci, va, n = matrix(name)
    m = rp.size - 1
    x = gen.make_x(n, 94)
    f = hip.csr_to_format(rp, ci, va, m, n, ci.size, kind=kind, sigma=256, lanes_per_row=T)
    lens = np.diff(rp)
    if kind == "sell_c_s" and m > 0:
        np.testing.assert_array_equal(f.export(0, np.int32), pyoracle.sellcs_convert(rp, ci, va, 64, 256)["order"])
    widths = f.export(1, np.int32)
    sptr = f.export(2, np.int64)
    np.testing.assert_array_equal(np.diff(sptr), ((widths + T - 1) // T * T) * 64)
    vals = f.export(4, np.float64)
    assert np.count_nonzero(vals) == np.count_nonzero(va)
    y = run(f, x, m)
    assert pyoracle.normwise_rel_err(y, pyoracle.csr_naive(rp, ci, va, x)) <= TOL
    assert (y[lens == 0] == 0).all()
    check_dd(y, rp, ci, va, x, "%s T%d/%s" % (kind, T, name))


@pytest.mark.parametrize("name", ALL)
@pytest.mark.parametrize("unroll", [1, 4, 8])
def test_sell_bit_exact(name, unroll):
    rp, ci, va, n = matrix(name)
    m = rp.size - 1
    print("case", name, unroll, flush=True)
    x = gen.make_x(n, 95)
    f = hip.csr_to_format(rp, ci, va, m, n, ci.size, kind="sell", unroll=unroll)
    ref = pyoracle.sell_convert(rp, ci, va, 64, T=1)
    np.testing.assert_array_equal(f.export(2, np.int64), ref["cluster_ptr"])
    np.testing.assert_array_equal(f.export(3, np.int32), ref["ja"])
    np.testing.assert_array_equal(f.export(4, np.float64), ref["a"])
    y = run(f, x, m)
    np.testing.assert_array_equal(y, pyoracle.sell_spmv(ref, 64, x, m))


def banded(m, half, seed):
    """Rows of 0..40 columns inside [r - half, r + half]."""
    rng = np.random.default_rng(seed)
    rows = []
    for r in range(m):
        lo, hi = max(0, r - half), min(m, r + half + 1)
        k = min(hi - lo, int(rng.integers(0, 41)))
        rows.append(np.sort(rng.choice(np.arange(lo, hi), size=k, replace=False)))
    rp = np.zeros(m + 1, np.int32)
    rp[1:] = np.cumsum([len(c) for c in rows])
    ci = np.concatenate(rows).astype(np.int32)
    return rp, ci, rng.uniform(-1, 1, ci.size), m


XW_CASES = ["banded_300", "banded_2500", "block_edges", "ragged_4097", "all_empty", "one_row_empty", "cfg1"]


def xw_matrix(name):
    if name == "banded_300":
        return banded(3001, 300, 11)
    if name == "banded_2500":
        return banded(9000, 2500, 12)
    return matrix(name)


@pytest.mark.parametrize("name", XW_CASES)
@pytest.mark.parametrize("kind,sigma,T,cap", [("sell_c_s", 256, 1, 1), ("sell_c_s", 4096, 1, 16384),
                                              ("sell_c_s", 64, 4, 1), ("sell", 1, 1, 8192), ("sell", 1, 2, 1)])
@pytest.mark.parametrize("pipeline", [0, 1])
def test_sell_x_window_and_pipeline(name, kind, sigma, T, cap, pipeline):
    """x read from a per-workgroup LDS window (columns outside it, i.e. padding
    entries, from global memory) and the software-pipelined loads keep every
    FMA in entry order: y is bit-identical to the same layout without them.
    The windows are checked against their definition (min / max column of the
    workgroup's 256 stored rows, staged only when the span fits)."""
    rp, ci, va, n = xw_matrix(name)
    m = rp.size - 1
    print("case", name, kind, sigma, T, cap, pipeline, flush=True)
    x = gen.make_x(n, 93)
    kw = dict(lanes_per_row=T, unroll=4)
    if kind == "sell_c_s":
        kw["sigma"] = sigma
    plain = hip.csr_to_format(rp, ci, va, m, n, ci.size, kind=kind, **kw)
    f = hip.csr_to_format(rp, ci, va, m, n, ci.size, kind=kind, x_window=cap, pipeline=pipeline, **kw)
    y0 = run(plain, x, m)
    y = run(f, x, m)
    np.testing.assert_array_equal(y, y0)
    if m == 0:
        return
    win = f.export(9, np.int32).reshape(-1, 2)
    order = f.export(0, np.int32) if kind == "sell_c_s" else np.arange(m, dtype=np.int32)
    limit = 2048 if cap == 1 else cap
    staged = 0
    for b in range(win.shape[0]):
        rows = order[b * 256:(b + 1) * 256]
        cols = np.concatenate([ci[rp[r]:rp[r + 1]] for r in rows]) if rows.size else np.zeros(0, np.int32)
        if cols.size and cols.max() - cols.min() + 1 <= limit:
            assert tuple(win[b]) == (cols.min(), cols.max() - cols.min() + 1)
            staged += 1
        else:
            assert tuple(win[b]) == (0, 0)
    if name.startswith("banded") and limit >= 2 * int(name.split("_")[1]) + 256:
        assert staged > 0  # the band plus a workgroup's rows fits: windows must be staged
    check_dd(y, rp, ci, va, x, "%s xw%d pipe%d/%s" % (kind, cap, pipeline, name))


@pytest.mark.parametrize("name", ["cfg1", "ragged_4097", "huge_rows", "block_edges", "one_row_empty", "all_empty",
                                  "dense_rows"])
@pytest.mark.parametrize("kind,sigma,T", [("sell_c_s", 64, 1), ("sell_c_s", 256, 4), ("sell_c_s", 4096, 1),
                                          ("sell_c_s", 4096, 8), ("sell", 1, 1), ("sell", 1, 4)])
@pytest.mark.parametrize("y_mode", [1, 2, 3])
def test_sell_y_modes(name, kind, sigma, T, y_mode):
    """Direct stores, LDS staging and the sorted buffer + permutation pass only
    change where each result is stored: y is bit-identical to the default."""
    rp, ci, va, n = matrix(name)
    m = rp.size - 1
    print("case", name, kind, sigma, T, y_mode, flush=True)
    x = gen.make_x(n, 92)
    kw = dict(lanes_per_row=T, unroll=4)
    if kind == "sell_c_s":
        kw["sigma"] = sigma
    if y_mode == 2 and kind == "sell_c_s" and 256 % sigma:
        with pytest.raises(hip.SpmvHipError):
            hip.csr_to_format(rp, ci, va, m, n, ci.size, kind=kind, y_mode=y_mode, **kw)
        return
    y0 = run(hip.csr_to_format(rp, ci, va, m, n, ci.size, kind=kind, **kw), x, m)
    y = run(hip.csr_to_format(rp, ci, va, m, n, ci.size, kind=kind, y_mode=y_mode, **kw), x, m)
    np.testing.assert_array_equal(y, y0)
    assert (y[np.diff(rp) == 0] == 0).all()


def test_bad_parameters_are_errors():
    rp, ci, va, n = matrix("test_txt")
    m = rp.size - 1
    for kind, kw in [("sell_c_s", dict(sigma=20000)), ("sell", dict(C=32)),
                     ("csr_vec", dict(block_nnz=8)), ("csr_vec", dict(lanes_per_row=3)),
                     ("sell_c_s", dict(x_window=20000))]:
        with pytest.raises(hip.SpmvHipError) as e:
            hip.csr_to_format(rp, ci, va, m, n, ci.size, kind=kind, **kw)
        assert e.value.code == -1


def test_repeated_calls_and_many_handles():
    """Handles are independent (no file-scope state); repeated calls are stable."""
    fmts = []
    ys = []
    for name in ("test_txt", "batch_03"):
        rp, ci, va, n = matrix(name)
        m = rp.size - 1
        x = gen.make_x(n, 1)
        for kind in ("csr_vec", "sell_c_s", "sell"):
            f = hip.csr_to_format(rp, ci, va, m, n, ci.size, kind=kind)
            fmts.append((f, x, m))
    for f, x, m in fmts:
        y1 = np.empty(m)
        f.spmv(x, y1)
        ys.append(y1)
    for (f, x, m), y1 in zip(fmts, ys):
        for _ in range(3):
            y2 = np.empty(m)
            f.spmv(x, y2)
            np.testing.assert_array_equal(y1, y2)


def test_bound_host_xy_mode():
    """GPU_KERNEL=1 driver mode: x/y bound once, y synchronised on request."""
    rp, ci, va, n = matrix("cfg1")
    m = rp.size - 1
    x = gen.make_x(n, 3)
    y = np.full(m, 1.0)
    f = hip.csr_to_format(rp, ci, va, m, n, ci.size, kind="csr_vec")
    hip._check(hip._lib.spmvhip_bind_host_xy(f._h, x.ctypes.data, y.ctypes.data), "bind")
    for _ in range(5):
        f.spmv(x, y)
    assert (y == 1.0).all()  # still on the device
    hip._check(hip._lib.spmvhip_sync_host_y(f._h), "sync")
    assert pyoracle.normwise_rel_err(y, pyoracle.csr_naive(rp, ci, va, x)) <= TOL
