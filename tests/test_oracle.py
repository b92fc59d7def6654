"""Pinning of the CPU oracle: against the reference's own conversion output
(golden fixtures), brute force on tiny inputs, closed forms and exact
arithmetic -- never against a retyped copy of its own formula."""
import hashlib
import json
import os
from fractions import Fraction

import numpy as np
import pytest

from spmvdare import gen
import pyoracle

HERE = os.path.dirname(os.path.abspath(__file__))
GEN = json.load(open(os.path.join(HERE, "golden", "generator.json")))
SELL = json.load(open(os.path.join(HERE, "golden", "sellcs.json")))
SSORT = json.load(open(os.path.join(HERE, "golden", "sell_sorted.json")))
SELLREF = json.load(open(os.path.join(HERE, "golden", "sell.json")))


def digest(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()[:32]


def random_csr(rng, m, n, max_len, empty_frac=0.2):
    lens = rng.integers(0, max_len + 1, size=m)
    lens[rng.random(m) < empty_frac] = 0
    lens = np.minimum(lens, n)
    rp = np.zeros(m + 1, dtype=np.int32)
    rp[1:] = np.cumsum(lens)
    ci = np.concatenate([np.sort(rng.choice(n, size=k, replace=False)) for k in lens]).astype(np.int32) \
        if lens.sum() else np.zeros(0, dtype=np.int32)
    va = rng.uniform(-1, 1, size=int(lens.sum()))
    return rp, ci, va


# ------------------------------------------------------------ SELL-C-sigma

@pytest.mark.parametrize("key", sorted(SELL))
def test_sellcs_oracle_matches_reference(key):
    f = SELL[key]
    a = gen.generate(gen.GenParams.parse(f["line"]), stats=False, map_values=False)
    conv = pyoracle.sellcs_convert(a.row_ptr, a.col_ind, a.values, f["C"], f["sigma"], width_rule=0)
    assert conv["truncated"] == 0
    assert int(conv["slice_ptr"][-1]) == f["total"]
    for k, fk in (("order", "order"), ("widths", "widths"), ("slice_ptr", "slice_ptr"),
                  ("cols", "cols"), ("vals", "vals")):
        assert digest(conv[k]) == f[fk], k


def test_sellcs_tiny_full_arrays(golden_dir):
    csr = np.load(os.path.join(golden_dir, "tiny_csr.npz"))
    ref = np.load(os.path.join(golden_dir, "tiny_sellcs.npz"))
    conv = pyoracle.sellcs_convert(csr["row_ptr"], csr["col_ind"], csr["values"], 64, 256)
    for k in ("order", "widths", "slice_ptr", "cols", "vals"):
        np.testing.assert_array_equal(conv[k], ref[k])


@pytest.mark.parametrize("sigma", [1, 3, 64, 100, 256])
def test_sellcs_order_brute_force(sigma):
    rng = np.random.default_rng(sigma)
    rp, ci, va = random_csr(rng, 517, 300, 12)
    lens = np.diff(rp)
    conv = pyoracle.sellcs_convert(rp, ci, va, 64, sigma, width_rule=1)
    expect = []
    for w0 in range(0, 517, sigma):
        ids = list(range(w0, min(517, w0 + sigma)))
        expect += sorted(ids, key=lambda i: (-lens[i], i))
    np.testing.assert_array_equal(conv["order"], expect)
    # width rule 1: slice width = longest row of the slice
    for s in range(conv["widths"].size):
        rows = conv["order"][s * 64:(s + 1) * 64]
        assert conv["widths"][s] == lens[rows].max()
    assert conv["truncated"] == 0


def test_sellcs_spmv_oracle_against_dense():
    rng = np.random.default_rng(3)
    rp, ci, va = random_csr(rng, 300, 200, 20)
    x = rng.uniform(-1, 1, 200)
    conv = pyoracle.sellcs_convert(rp, ci, va, 64, 128)
    y = pyoracle.sellcs_spmv(conv, 64, x)
    dense = np.zeros((300, 200))
    for i in range(300):
        dense[i, ci[rp[i]:rp[i + 1]]] = va[rp[i]:rp[i + 1]]
    np.testing.assert_allclose(y, dense @ x, rtol=1e-12, atol=1e-13)


# ------------------------------------------------------------ SELL_SORTED

@pytest.mark.parametrize("key", sorted(SSORT))
def test_sellsorted_oracle_matches_reference(key):
    """Thread windows, permutation, cluster sizes, ja, a and y of the
    reference's own SELL_SORTED build (sell_sorted.cpp, 4-row clusters)."""
    f = SSORT[key]
    a = gen.generate(gen.GenParams.parse(f["line"]), stats=False, map_values=False)
    win = pyoracle.sellsorted_windows(a.row_ptr, f["VL"], f["T"])
    assert win.tolist() == f["windows"]
    conv = pyoracle.sellsorted_convert(a.row_ptr, a.col_ind, a.values, f["VL"], win)
    assert int(conv["cluster_ptr"][-1]) == f["nnz_ext"]
    for k in ("perm", "cluster_ptr", "ja", "a"):
        assert digest(conv[k]) == f[k], k
    x = ((np.arange(a.n) % 13) + 1) / 8.0 - 1.0
    assert digest(pyoracle.sellsorted_spmv(conv, f["VL"], x)) == f["y"]


@pytest.mark.parametrize("VL,bounds", [(4, [0, 12, 40, 101]), (64, [0, 128, 517]), (8, [0, 517])])
def test_sellsorted_order_brute_force(VL, bounds):
    rng = np.random.default_rng(VL)
    rp, ci, va = random_csr(rng, 517, 300, 12)
    if bounds[-1] != 517:
        rp = rp[:bounds[-1] + 1]
    m = rp.size - 1
    lens = np.diff(rp)
    conv = pyoracle.sellsorted_convert(rp, ci, va, VL, bounds)
    order = []
    for w0, w1 in zip(bounds[:-1], bounds[1:]):
        order += sorted(range(w0, w1), key=lambda i: (lens[i], i))
    np.testing.assert_array_equal(conv["order"], order)
    for c in range((m + VL - 1) // VL):
        rows = conv["order"][c * VL:(c + 1) * VL]
        assert conv["cluster_ptr"][c + 1] - conv["cluster_ptr"][c] == VL * lens[rows].max()
    x = rng.uniform(-1, 1, 300)
    dense = np.zeros((m, 300))
    for i in range(m):
        dense[i, ci[rp[i]:rp[i + 1]]] = va[rp[i]:rp[i + 1]]
    np.testing.assert_allclose(pyoracle.sellsorted_spmv(conv, VL, x), dense @ x, rtol=1e-12, atol=1e-13)


# ------------------------------------------------------------ SELL

@pytest.mark.parametrize("key", sorted(SELLREF))
def test_sell_oracle_matches_reference(key):
    """Cluster sizes, ja (with the per-thread padding column carry), a and y
    of the reference's own SELL build (sell.cpp, 4-row clusters, T threads)."""
    f = SELLREF[key]
    a = gen.generate(gen.GenParams.parse(f["line"]), stats=False, map_values=False)
    conv = pyoracle.sell_convert(a.row_ptr, a.col_ind, a.values, f["VL"], T=f["T"])
    assert int(conv["cluster_ptr"][-1]) == f["nnz_ext"]
    for k in ("cluster_ptr", "ja", "a"):
        assert digest(conv[k]) == f[k], k
    x = ((np.arange(a.n) % 13) + 1) / 8.0 - 1.0
    assert digest(pyoracle.sell_spmv(conv, f["VL"], x, a.m)) == f["y"]


def test_sell_padding_rule_closed_form():
    """sell.cpp:165-192: padding takes value 0 and the column of the previous
    entry written by the same thread (the row's last column, carried across
    empty rows; 0 before the thread's first entry)."""
    # rows: [1,5], [], [7], [], [], [2,3,9]  -> VL = 4, clusters {0..3}, {4,5}
    rp = np.array([0, 2, 2, 3, 3, 3, 6], dtype=np.int32)
    ci = np.array([1, 5, 7, 2, 3, 9], dtype=np.int32)
    va = np.arange(1, 7, dtype=np.float64)
    conv = pyoracle.sell_convert(rp, ci, va, 4, T=1)
    # cluster 0: width 2; row-major rows: [1,5] [5,5](carry) [7,7] [7,7]
    c0 = np.array([[1, 5], [5, 5], [7, 7], [7, 7]]).T.reshape(-1)
    # cluster 1: width 3; rows: [7,7,7](empty row 4 carries 7) [2,3,9] [9,9,9] [9,9,9]
    c1 = np.array([[7, 7, 7], [2, 3, 9], [9, 9, 9], [9, 9, 9]]).T.reshape(-1)
    np.testing.assert_array_equal(conv["cluster_ptr"], [0, 8, 20])
    np.testing.assert_array_equal(conv["ja"], np.concatenate([c0, c1]))
    assert (conv["a"][conv["ja"] == 0] == 0).all()
    x = np.arange(10, dtype=np.float64) + 1
    y = pyoracle.sell_spmv(conv, 4, x, 6)
    np.testing.assert_array_equal(y, [1 * 2 + 2 * 6, 0, 3 * 8, 0, 0, 4 * 3 + 5 * 4 + 6 * 10])


def test_sell_thread_partition_resets_carry():
    """With two threads the second thread's carry restarts at column 0."""
    rp = np.array([0, 2, 2, 3, 3, 3, 6, 6, 8], dtype=np.int32)
    ci = np.array([1, 5, 7, 2, 3, 9, 4, 6], dtype=np.int32)
    va = np.ones(8)
    one = pyoracle.sell_convert(rp, ci, va, 2, T=1)
    two = pyoracle.sell_convert(rp, ci, va, 2, T=2)
    s, e = pyoracle.partition_prefix_sums(2, 1, one["cluster_ptr"][:-1], int(one["cluster_ptr"][-1]))
    assert s > 0
    # rows of the second thread's first cluster: differs only where it starts with an empty row
    assert not np.array_equal(one["ja"], two["ja"]) or np.diff(rp)[2 * s] > 0


# ------------------------------------------------------------ csr_naive / dd

def test_csr_naive_closed_form():
    rp = np.array([0, 3, 3, 5], dtype=np.int32)
    ci = np.array([0, 2, 3, 1, 3], dtype=np.int32)
    va = np.array([0.5, -2.0, 4.0, 3.0, 0.25])
    x = np.array([2.0, -1.0, 0.5, 8.0])
    y = pyoracle.csr_naive(rp, ci, va, x)
    np.testing.assert_array_equal(y, [0.5 * 2 - 2 * 0.5 + 4 * 8, 0.0, -3.0 + 2.0])


def test_csr_naive_rounds_each_step():
    """A left-to-right sum with separate rounding: (1e16 + 1) - 1e16 == 0, the
    fused or reordered sums would give 1."""
    rp = np.array([0, 3], dtype=np.int32)
    ci = np.array([0, 1, 2], dtype=np.int32)
    va = np.array([1.0, 1.0, -1.0])
    x = np.array([1e16, 1.0, 1e16])
    assert pyoracle.csr_naive(rp, ci, va, x)[0] == 0.0
    hi, lo = pyoracle.csr_dd(rp, ci, va, x)
    assert hi[0] + lo[0] == 1.0


def test_csr_dd_against_exact_rationals():
    rng = np.random.default_rng(11)
    rp, ci, va = random_csr(rng, 40, 60, 30, empty_frac=0.1)
    x = rng.uniform(-1, 1, 60) * 10.0 ** rng.integers(-8, 8, 60)
    hi, lo = pyoracle.csr_dd(rp, ci, va, x)
    for i in range(40):
        exact = sum((Fraction(va[j]) * Fraction(x[ci[j]]) for j in range(rp[i], rp[i + 1])), Fraction(0))
        got = Fraction(hi[i]) + Fraction(lo[i])
        scale = sum((abs(Fraction(va[j]) * Fraction(x[ci[j]])) for j in range(rp[i], rp[i + 1])), Fraction(0))
        assert abs(got - exact) <= scale * Fraction(1, 2 ** 100)


def test_csr_naive_omp_matches_serial():
    a = gen.generate(gen.GenParams.parse(GEN["batch_04"]["line"]))
    x = gen.make_x(a.n, 5)
    y1 = pyoracle.csr_naive(a.row_ptr, a.col_ind, a.values, x)
    y2, t = pyoracle.csr_naive_omp(a.row_ptr, a.col_ind, a.values, x, threads=4)
    assert t == 4
    np.testing.assert_array_equal(y1, y2)


# ------------------------------------------------------------ partitioners

def _closest(A, target):
    """Linear-scan statement of the closest-prefix search on a non-decreasing
    array: out-of-range targets clamp to the ends; otherwise take the first
    index whose value reaches the target, and step back one when the
    previous value is strictly closer (ties go to the later index)."""
    if target < A[0]:
        return 0
    if target > A[-1]:
        return len(A) - 1
    e = next(i for i, a in enumerate(A) if a >= target)
    if A[e] == target or e == 0:
        return e
    return e - 1 if target - A[e - 1] < A[e] - target else e


def test_partition_prefix_sums_brute_force():
    rng = np.random.default_rng(5)
    for trial in range(200):
        N = int(rng.integers(1, 40))
        w = rng.integers(0, 20, size=N)
        sums = np.concatenate([[0], np.cumsum(w)[:-1]]).astype(np.int64)
        total = int(w.sum())
        W = int(rng.integers(1, 9))
        prev_e = 0
        for k in range(W):
            s, e = pyoracle.partition_prefix_sums(W, k, sums, total)
            if k == 0:
                assert s == 0
            else:
                assert s == prev_e
            if k == W - 1:
                assert e == N
            else:
                target = total * (k + 1) // W
                assert e == _closest(list(sums), target)
            prev_e = e


def test_partition_iterations_brute_force():
    for length in range(0, 30):
        for W in range(1, 9):
            covered = []
            for k in range(W):
                s, e = pyoracle.partition_iterations(W, k, 0, length)
                covered += list(range(s, e))
            assert covered == list(range(length))
