"""Artificial matrix generator (host, native) and benchmark input definitions.

Mirrors artificial_matrix_generation() (artificial-matrix-generator/
artificial_matrix_generation.h:90) with the argument order of the reference's
command line (artificial-matrix-generator/artificial_matrix.c:40-50 and
src/bench.cpp:562-585):
    nr_rows nr_cols avg_nnz_per_row std_nnz_per_row distribution placement
    bw skew avg_num_neighbours cross_row_similarity seed
"""
import ctypes
import hashlib
import os
from dataclasses import dataclass, fields

import numpy as np

from . import _native

_lib = _native.load("spmvgen")


class _Stats(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in (
        "avg_nnz_per_row", "std_nnz_per_row", "min_nnz_per_row", "max_nnz_per_row",
        "skew", "avg_bw", "std_bw", "avg_bw_scaled", "avg_sc", "avg_num_neighbours",
        "cross_row_similarity", "avg_clustering", "avg_ngroups", "bw_corrected")]


class _CSR(ctypes.Structure):
    _fields_ = [("m", ctypes.c_long), ("n", ctypes.c_long), ("nnz", ctypes.c_long),
                ("row_ptr", ctypes.POINTER(ctypes.c_int32)),
                ("col_ind", ctypes.POINTER(ctypes.c_int32)),
                ("values", ctypes.POINTER(ctypes.c_double)),
                ("stats", _Stats)]


_lib.spmvgen_generate.argtypes = [
    ctypes.c_long, ctypes.c_long, ctypes.c_double, ctypes.c_double, ctypes.c_char_p,
    ctypes.c_uint, ctypes.c_char_p, ctypes.c_double, ctypes.c_double, ctypes.c_double,
    ctypes.c_double, ctypes.c_int, ctypes.POINTER(_CSR)]
_lib.spmvgen_generate.restype = ctypes.c_int
_lib.spmvgen_free.argtypes = [ctypes.POINTER(_CSR)]
_lib.spmvgen_free.restype = None
_lib.spmvgen_last_error.restype = ctypes.c_char_p
_lib.spmvgen_map_values.argtypes = [ctypes.c_void_p, ctypes.c_long]
_lib.spmvgen_map_values.restype = None
_lib.spmvgen_fill_x.argtypes = [ctypes.c_void_p, ctypes.c_long, ctypes.c_uint]
_lib.spmvgen_fill_x.restype = ctypes.c_int
_lib.spmvgen_degrees.argtypes = [
    ctypes.c_long, ctypes.c_long, ctypes.c_double, ctypes.c_double, ctypes.c_char_p,
    ctypes.c_uint, ctypes.c_double, ctypes.c_void_p, ctypes.POINTER(ctypes.c_long)]
_lib.spmvgen_degrees.restype = ctypes.c_int
_lib.spmvgen_compute_stats.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_long,
                                       ctypes.c_long, ctypes.POINTER(_Stats)]
_lib.spmvgen_compute_stats.restype = ctypes.c_int
_lib.spmvgen_read_mtx.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.POINTER(_CSR), ctypes.POINTER(ctypes.c_int)]
_lib.spmvgen_read_mtx.restype = ctypes.c_int


class GenError(RuntimeError):
    pass


@dataclass(frozen=True)
class GenParams:
    """The 11 generator descriptors, in the reference's command-line order."""
    nr_rows: int
    nr_cols: int
    avg_nnz_per_row: float
    std_nnz_per_row: float
    distribution: str
    placement: str
    bw: float
    skew: float
    avg_num_neighbours: float
    cross_row_similarity: float
    seed: int

    @classmethod
    def parse(cls, line):
        """One parameter line, e.g. matrix_generation_parameters/test.txt:1."""
        t = line.split()
        if len(t) < 11:
            raise ValueError("expected 11 generator arguments, got %d" % len(t))
        return cls(int(t[0]), int(t[1]), float(t[2]), float(t[3]), t[4], t[5],
                   float(t[6]), float(t[7]), float(t[8]), float(t[9]), int(t[10]))

    def line(self):
        return " ".join(str(getattr(self, f.name)) for f in fields(self))

    def descriptors(self):
        """The 10 numeric generator descriptors used as selector features."""
        return [float(self.nr_rows), float(self.nr_cols), self.avg_nnz_per_row,
                self.std_nnz_per_row, 1.0 if self.distribution.startswith("g") else 0.0,
                1.0 if self.placement.startswith("d") else 0.0, self.bw, self.skew,
                self.avg_num_neighbours, self.cross_row_similarity]


@dataclass
class CSRMatrix:
    m: int
    n: int
    row_ptr: np.ndarray   # int32, m + 1
    col_ind: np.ndarray   # int32, nnz
    values: np.ndarray    # float64, nnz
    stats: dict

    @property
    def nnz(self):
        return int(self.row_ptr[-1])


def generate(p, stats=True, map_values=True):
    """Generate one matrix. With map_values the values are 2 v - 1 in [-1, 1)
    (exact); otherwise the reference's raw values in [0, 1)."""
    c = _CSR()
    rc = _lib.spmvgen_generate(p.nr_rows, p.nr_cols, p.avg_nnz_per_row, p.std_nnz_per_row,
                               p.distribution.encode(), p.seed & 0xFFFFFFFF,
                               p.placement.encode(), p.bw, p.skew, p.avg_num_neighbours,
                               p.cross_row_similarity, 1 if stats else 0, ctypes.byref(c))
    if rc != 0:
        raise GenError("spmvgen_generate(%s): %s" % (p.line(), _lib.spmvgen_last_error().decode()))
    try:
        m, nnz = c.m, c.nnz
        rp = np.ctypeslib.as_array(c.row_ptr, shape=(m + 1,)).copy()
        ci = np.ctypeslib.as_array(c.col_ind, shape=(max(nnz, 1),))[:nnz].copy()
        va = np.ctypeslib.as_array(c.values, shape=(max(nnz, 1),))[:nnz].copy()
        st = {f[0]: getattr(c.stats, f[0]) for f in _Stats._fields_}
    finally:
        _lib.spmvgen_free(ctypes.byref(c))
    if map_values:
        map_values_inplace(va)
    return CSRMatrix(m, c.n, rp, ci, va, st)


def generate_cached(p, cache_dir, stats=False, map_values=True):
    """generate(p) through an on-disk cache keyed by the 11 parameters (and
    the value mapping): the first call stores row_ptr / col_ind / values as
    .npy files under cache_dir, later calls memory-map them. Without stats
    (the cache does not keep them)."""
    if stats:
        raise ValueError("generate_cached keeps no statistics; use generate(p, stats=True)")
    key = hashlib.sha256(("%s|%d" % (p.line(), 1 if map_values else 0)).encode()).hexdigest()[:32]
    d = os.path.join(cache_dir, key)
    names = ("row_ptr.npy", "col_ind.npy", "values.npy", "shape.npy")
    if all(os.path.exists(os.path.join(d, f)) for f in names):
        m, n = (int(v) for v in np.load(os.path.join(d, "shape.npy")))
        rp, ci, va = (np.load(os.path.join(d, f), mmap_mode="r") for f in names[:3])
        return CSRMatrix(m, n, rp, ci, va, {})
    a = generate(p, stats=False, map_values=map_values)
    tmp = d + ".tmp%d" % os.getpid()
    os.makedirs(tmp, exist_ok=True)
    for f, arr in zip(names, (a.row_ptr, a.col_ind, a.values, np.array([a.m, a.n], dtype=np.int64))):
        np.save(os.path.join(tmp, f), arr)
    try:
        os.rename(tmp, d)
    except OSError:  # another process stored it first
        for f in names:
            os.remove(os.path.join(tmp, f))
        os.rmdir(tmp)
    return a


def map_values_inplace(values):
    assert values.dtype == np.float64 and values.flags.c_contiguous
    _lib.spmvgen_map_values(values.ctypes.data, values.size)


def make_x(n, seed):
    """x_i = 2 u_i - 1, u from one glibc random_r stream seeded `seed`."""
    x = np.empty(n, dtype=np.float64)
    rc = _lib.spmvgen_fill_x(x.ctypes.data, n, seed & 0xFFFFFFFF)
    if rc != 0:
        raise GenError(_lib.spmvgen_last_error().decode())
    return x


def degrees(p):
    d = np.empty(p.nr_rows, dtype=np.int32)
    nnz = ctypes.c_long()
    rc = _lib.spmvgen_degrees(p.nr_rows, p.nr_cols, p.avg_nnz_per_row, p.std_nnz_per_row,
                              p.distribution.encode(), p.seed & 0xFFFFFFFF, p.skew,
                              d.ctypes.data, ctypes.byref(nnz))
    if rc != 0:
        raise GenError(_lib.spmvgen_last_error().decode())
    return d, nnz.value


def structure_stats(row_ptr, col_ind, n):
    s = _Stats()
    rp = np.ascontiguousarray(row_ptr, dtype=np.int32)
    ci = np.ascontiguousarray(col_ind, dtype=np.int32)
    rc = _lib.spmvgen_compute_stats(rp.ctypes.data, ci.ctypes.data, rp.size - 1, n, ctypes.byref(s))
    if rc != 0:
        raise GenError(_lib.spmvgen_last_error().decode())
    return {f[0]: getattr(s, f[0]) for f in _Stats._fields_}


class MtxError(GenError):
    pass


def read_mtx(path, expand_symmetry=True):
    """A Matrix Market file as CSR (spmvgen_read_mtx: the reference's
    import_file semantics). Returns (CSRMatrix, symmetric flag of the file)."""
    c = _CSR()
    sym = ctypes.c_int()
    rc = _lib.spmvgen_read_mtx(os.fsencode(path), 1 if expand_symmetry else 0, ctypes.byref(c), ctypes.byref(sym))
    if rc != 0:
        raise MtxError("spmvgen_read_mtx(%s): %s" % (path, _lib.spmvgen_last_error().decode()))
    try:
        m, nnz = c.m, c.nnz
        rp = np.ctypeslib.as_array(c.row_ptr, shape=(m + 1,)).copy()
        ci = np.ctypeslib.as_array(c.col_ind, shape=(max(nnz, 1),))[:nnz].copy()
        va = np.ctypeslib.as_array(c.values, shape=(max(nnz, 1),))[:nnz].copy()
    finally:
        _lib.spmvgen_free(ctypes.byref(c))
    return CSRMatrix(m, c.n, rp, ci, va, {}), bool(sym.value)
