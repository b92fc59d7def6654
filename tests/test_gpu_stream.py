"""End-to-end streaming (spmvdare/stream.py) on the first 16 matrices of the
held-out stream (master seed 55: rows 1e4..4e6, avg 2..128 nnz/row, gamma
rows, skew up to 1000): the committed selector chooses every matrix's format
inside the batched launch, with its fallback to exhaustive tuning, and every
batched y is checked against the oracle's csr_naive at 1e-13 normwise."""
import os

import numpy as np
import pytest

from spmvdare import gen, hip, stream, tuning
import pyoracle

pytestmark = [pytest.mark.gpu, pytest.mark.timeout(900)]
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MODEL = os.path.join(ROOT, "spmv-dare_amd", "spmvdare", "models", "selector.json")
RANGES = tuning.STREAM_RANGES


@pytest.fixture(scope="module")
def selector():
    hip.init(0)
    return tuning.Selector.load(MODEL)


def test_stream_first_16_matrices(selector):
    params = tuning.sample_params(55, 64, **RANGES)
    r = stream.run_stream(params, 16, 8, selector, tuned=2, fallback_spread=0.3, steps=3, keep_y=True,
                          log=lambda s: print(s, flush=True))
    print({k: v for k, v in r.items() if k != "y"}, flush=True)
    assert r["matrices"] == 16 and len(r["y"]) == 16
    assert r["max_normwise_err_vs_csr_naive"] <= 1e-13
    assert sum(r["chosen_variants"].values()) == 16
    assert r["regret"]["matrices"] == 2
    for p, yb in r["y"]:
        a = gen.generate(p, stats=False)
        x = gen.make_x(a.n, p.seed + 1)
        y_ref, _ = pyoracle.csr_naive_omp(a.row_ptr, a.col_ind, a.values, x)
        err = pyoracle.normwise_rel_err(yb, y_ref)
        print("stream member rows=%d nnz=%d: normwise vs csr_naive %.3e" % (a.m, a.nnz, err), flush=True)
        assert err <= 1e-13


def test_stream_fallback_is_exhaustive(selector):
    """With a negative spread limit every choice falls back to exhaustive
    tuning, so the policy's regret against that same tuning is 0."""
    params = tuning.sample_params(21, 8)   # small matrices of the batched workload's ranges
    r = stream.run_stream(params, 3, 3, selector, tuned=3, fallback_spread=-0.5, steps=2)
    assert r["fallbacks"] == 3
    assert r["regret"]["policy_max"] == 0.0
    assert r["max_normwise_err_vs_csr_naive"] <= 1e-13


def test_stream_raises_when_host_generation_fails(selector, monkeypatch):
    """A failure of the host thread that generates the next chunk reaches the
    caller (the stream does not end early as if it were done)."""
    params = tuning.sample_params(21, 8)   # small matrices of the batched workload's ranges
    real = stream.gen.generate
    calls = {"n": 0}

    def failing(p, *a, **kw):
        calls["n"] += 1
        if calls["n"] > 2:  # the first chunk (2 matrices) generates, the next one fails
            raise MemoryError("host generation failed")
        return real(p, *a, **kw)

    monkeypatch.setattr(stream.gen, "generate", failing)
    with pytest.raises(RuntimeError) as e:
        stream.run_stream(params, 4, 2, selector, steps=1)
    assert isinstance(e.value.__cause__, MemoryError)
