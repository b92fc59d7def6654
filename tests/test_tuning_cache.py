"""Autotune cache and matrix keys (host logic; the timing itself is covered
by the GPU tests)."""
import json

import numpy as np

from spmvdare import gen, tuning


def test_autotune_cache_persists_per_device(tmp_path):
    path = str(tmp_path / "tune.json")
    c = tuning.AutotuneCache(path, device="gfx950-a")
    assert c.get("k") == {}
    c.put("k", {"csr_vec(block_nnz=256,lanes_per_row=0)": 0.25})
    c2 = tuning.AutotuneCache(path, device="gfx950-a")
    assert c2.get("k") == {"csr_vec(block_nnz=256,lanes_per_row=0)": 0.25}
    assert tuning.AutotuneCache(path, device="other").get("k") == {}
    assert json.load(open(path))["gfx950-a"]["k"]
    # a damaged file starts an empty cache instead of failing
    open(path, "w").write("{")
    assert tuning.AutotuneCache(path, device="gfx950-a").get("k") == {}


def test_matrix_key():
    p = gen.GenParams.parse("1000 1000 5 4 normal random 0.3 10 1 0.2 14")
    a = gen.generate(p, stats=False)
    assert tuning.matrix_key(a, p) == "gen:" + p.line()
    k = tuning.matrix_key(a)
    assert k.startswith("sha256:") and k == tuning.matrix_key(a)
    b = gen.generate(p, stats=False)
    b.values[0] += 1.0
    assert tuning.matrix_key(b) != k
