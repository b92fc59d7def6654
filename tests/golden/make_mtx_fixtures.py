"""Writes the Matrix Market fixtures under tests/golden/mtx/ and records, in
tests/golden/mtx.json, what the reference's own import path reads from each
(oracle/_ref/mtx_dump: mtx_read + mtx_values_convert_to_real + coo_to_csr,
compiled from the reference sources by oracle/Makefile). Run in a container
where the reference is present:

    make -C oracle _ref/mtx_dump && python3 tests/golden/make_mtx_fixtures.py
"""
import json
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "spmv-dare_amd"))
from spmvdare import gen  # noqa: E402

OUT = os.path.join(HERE, "mtx")
DUMP = os.path.join(ROOT, "oracle", "_ref", "mtx_dump")


def write(name, banner, m, n, entries, comments=("% written by make_mtx_fixtures.py",), blank_lines=False):
    path = os.path.join(OUT, name + ".mtx")
    with open(path, "w") as f:
        if banner:
            f.write(banner + "\n")
        for c in comments:
            f.write(c + "\n")
        f.write("%d %d %d\n" % (m, n, len(entries)))
        for k, e in enumerate(entries):
            f.write(" ".join(e) + "\n")
            if blank_lines and k % 7 == 3:
                f.write("\n")
    return path


def main():
    os.makedirs(OUT, exist_ok=True)
    rng = np.random.default_rng(12)
    files = {}
    # general real: a generated matrix, entries in shuffled order
    a = gen.generate(gen.GenParams.parse("300 280 6 3 gamma random 0.4 4 1 0.2 5"), stats=False)
    rows = np.repeat(np.arange(a.m), np.diff(a.row_ptr))
    perm = rng.permutation(a.nnz)
    files["general_real"] = write("general_real", "%%MatrixMarket matrix coordinate real general", a.m, a.n,
                                  [("%d" % (rows[k] + 1), "%d" % (a.col_ind[k] + 1), "%.17g" % a.values[k])
                                   for k in perm], blank_lines=True)
    # symmetric real: the lower triangle of B + B^T (diagonal included)
    n = 200
    pairs = {}
    for _ in range(900):
        i, j = sorted(rng.integers(0, n, 2))[::-1]
        pairs[(int(i), int(j))] = float(rng.uniform(-1, 1))
    for i in range(0, n, 3):
        pairs[(i, i)] = float(rng.uniform(-1, 1))
    lower = sorted(pairs.items(), key=lambda kv: (kv[0][1], kv[0][0]))
    files["symmetric_real"] = write("symmetric_real", "%%MatrixMarket matrix coordinate real symmetric", n, n,
                                    [("%d" % (i + 1), "%d" % (j + 1), "%.17g" % v) for (i, j), v in lower])
    strict = [kv for kv in lower if kv[0][0] != kv[0][1]]
    files["skew_real"] = write("skew_real", "%%MatrixMarket matrix coordinate real skew-symmetric", n, n,
                               [("%d" % (i + 1), "%d" % (j + 1), "%.17g" % v) for (i, j), v in strict])
    files["symmetric_pattern"] = write("symmetric_pattern", "%%MatrixMarket matrix coordinate pattern symmetric",
                                       n, n, [("%d" % (i + 1), "%d" % (j + 1)) for (i, j), v in lower])
    files["general_integer"] = write("general_integer", "%%MatrixMarket matrix coordinate integer general", 50, 70,
                                     [("%d" % (rng.integers(1, 51)), "%d" % c, "%d" % rng.integers(-9, 10))
                                      for c in rng.permutation(70)[:60] + 1])
    cz = sorted({(int(rng.integers(1, 41)), int(rng.integers(1, 41))) for _ in range(150)})
    files["general_complex"] = write("general_complex", "%%MatrixMarket matrix coordinate complex general", 40, 40,
                                     [("%d" % r, "%d" % c, "%.17g" % rng.normal(), "%.17g" % rng.normal())
                                      for r, c in cz])
    hz = sorted({(int(max(r, c)), int(min(r, c))) for r, c in cz})
    files["hermitian_complex"] = write("hermitian_complex", "%%MatrixMarket matrix coordinate complex Hermitian",
                                       40, 40, [("%d" % r, "%d" % c, "%.17g" % rng.normal(),
                                                 "%.17g" % (0.0 if r == c else rng.normal())) for r, c in hz])
    files["no_banner"] = write("no_banner", None, 6, 5,
                               [("%d" % r, "%d" % c, "%.17g" % v) for r, c, v in
                                [(1, 1, 0.5), (6, 5, -2.25), (3, 2, 1e-300), (2, 4, 3.0), (6, 1, 7.5)]],
                               comments=())
    files["empty_rows"] = write("empty_rows", "%%MatrixMarket matrix coordinate real general", 10, 10,
                                [("4", "9", "1.5"), ("4", "2", "-1"), ("10", "10", "2")],
                                comments=("%", "% two comment lines, rows 1-3, 5-9 empty"))
    files["no_entries"] = write("no_entries", "%%MatrixMarket matrix coordinate real general", 3, 4, [])
    out = {}
    for name, path in sorted(files.items()):
        r = subprocess.run([DUMP, path], capture_output=True, text=True, check=True)
        m, n_, nnz, sym, h_rp, h_ci, h_va = r.stdout.split()
        out[name] = {"file": os.path.basename(path), "m": int(m), "n": int(n_), "nnz": int(nnz),
                     "symmetric": int(sym), "row_ptr": h_rp, "col_ind": h_ci, "values": h_va}
        print(name, out[name])
    json.dump(out, open(os.path.join(HERE, "mtx.json"), "w"), indent=1)


if __name__ == "__main__":
    main()
