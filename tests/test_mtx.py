"""Matrix Market import (spmvgen_read_mtx) against the reference's own import
path: for every fixture file, the CSR arrays hash equal to what the
reference's mtx_read + mtx_values_convert_to_real + coo_to_csr produce
(tests/golden/mtx.json, written by tests/golden/make_mtx_fixtures.py from the
compiled reference). Independent checks: a generated matrix written out
reads back exactly; symmetric expansion gives A = A^T, skew-symmetric
A = -A^T; malformed files are rejected with the reader's error code."""
import json
import os

import numpy as np
import pytest

from spmvdare import gen
import pyoracle

HERE = os.path.dirname(os.path.abspath(__file__))
MTX = json.load(open(os.path.join(HERE, "golden", "mtx.json")))


def path(name):
    return os.path.join(HERE, "golden", "mtx", MTX[name]["file"])


def fnv(a):
    return pyoracle.fnv1a(np.ascontiguousarray(a))


@pytest.mark.parametrize("name", sorted(MTX))
def test_matches_reference_import(name):
    a, sym = gen.read_mtx(path(name))
    ref = MTX[name]
    assert (a.m, a.n, a.nnz, int(sym)) == (ref["m"], ref["n"], ref["nnz"], ref["symmetric"])
    assert fnv(a.row_ptr) == ref["row_ptr"]
    assert fnv(a.col_ind) == ref["col_ind"]
    assert fnv(a.values) == ref["values"]


def dense(a):
    d = np.zeros((a.m, a.n))
    rows = np.repeat(np.arange(a.m), np.diff(a.row_ptr))
    np.add.at(d, (rows, a.col_ind), a.values)
    return d


def test_general_reads_back_the_written_matrix():
    src = gen.generate(gen.GenParams.parse("300 280 6 3 gamma random 0.4 4 1 0.2 5"), stats=False)
    a, sym = gen.read_mtx(path("general_real"))
    assert not sym
    np.testing.assert_array_equal(a.row_ptr, src.row_ptr)
    np.testing.assert_array_equal(a.col_ind, src.col_ind)
    np.testing.assert_array_equal(a.values, src.values)


def test_symmetries_expand():
    s = dense(gen.read_mtx(path("symmetric_real"))[0])
    np.testing.assert_array_equal(s, s.T)
    k = dense(gen.read_mtx(path("skew_real"))[0])
    np.testing.assert_array_equal(k, -k.T)
    p = gen.read_mtx(path("symmetric_pattern"))[0]
    assert (p.values == 1.0).all()
    lower, _ = gen.read_mtx(path("symmetric_real"), expand_symmetry=False)
    rows = np.repeat(np.arange(lower.m), np.diff(lower.row_ptr))
    assert (lower.col_ind <= rows).all()   # only the stored triangle


def test_columns_sorted_within_rows():
    for name in MTX:
        a = gen.read_mtx(path(name))[0]
        for r in range(a.m):
            c = a.col_ind[a.row_ptr[r]:a.row_ptr[r + 1]]
            assert (np.diff(c) >= 0).all()


@pytest.mark.parametrize("text", [
    "%%MatrixMarket matrix array real general\n2 2\n1\n2\n3\n4\n",          # dense: not read
    "%%MatrixMarket matrix coordinate real general\n3 3 2\n1 1 1.0\n",       # fewer entries than declared
    "%%MatrixMarket matrix coordinate real symmetric-ish\n1 1 1\n1 1 1\n",   # unknown symmetry word
    "%%MatrixMarket matrix coordinate real general\n2 2 1\n3 1 1.0\n",       # index outside the matrix
    "%%MatrixMarket vector coordinate real general\n2 2 1\n1 1 1.0\n",       # not a matrix
])
def test_malformed_files_rejected(tmp_path, text):
    f = tmp_path / "bad.mtx"
    f.write_text(text)
    with pytest.raises(gen.MtxError) as e:
        gen.read_mtx(str(f))
    assert "Matrix Market" in str(e.value)
