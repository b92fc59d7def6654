"""The committed selector model and its corpora index the current variant
list: the MLP's output slot i is tuning.VARIANTS[i], so a model trained on
another list would silently pick different kernels (CPU only)."""
import json
import os

import numpy as np
import pytest

from spmvdare import tuning

MODELS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "..", "spmv-dare_amd", "spmvdare", "models")
NAMES = [tuning.variant_name(i) for i in range(len(tuning.VARIANTS))]


def test_model_and_corpora_use_the_current_variants():
    d = json.load(open(os.path.join(MODELS, "selector.json")))
    assert d["variants"] == NAMES
    assert d["dims"] == [14, 64, 64, len(NAMES)]
    for f in ("corpus_train.npz", "corpus_test.npz", "stream_train.npz", "stream_test.npz"):
        c = np.load(os.path.join(MODELS, f), allow_pickle=False)
        assert list(c["variants"]) == NAMES
        assert c["times"].shape[1] == len(NAMES)


def test_load_rejects_another_variant_list(tmp_path):
    d = json.load(open(os.path.join(MODELS, "selector.json")))
    d["variants"] = list(reversed(d["variants"]))
    p = tmp_path / "stale.json"
    p.write_text(json.dumps(d))
    with pytest.raises(ValueError):
        tuning.Selector.load(str(p))


def test_stream_ranges_draws():
    """The streaming workload's draws: rows = cols log-uniform in [1e4, 4e6],
    avg 2..128, gamma rows only, skew up to 1000; the batched workload's
    draws stay within its own ranges."""
    ps = tuning.sample_params(56, 400, **tuning.WORKLOAD_RANGES["stream"])
    rows = np.array([p.nr_rows for p in ps])
    assert rows.min() >= 1e4 and rows.max() <= 4e6
    assert np.median(rows) > 2e5  # log-uniform over [1e4, 4e6]: median 2e5
    assert {p.distribution for p in ps} == {"gamma"}
    assert all(2 <= p.avg_nnz_per_row <= 128 and 1 <= p.skew <= 1000 for p in ps)
    assert all(p.nr_rows == p.nr_cols for p in ps)
    assert max(p.skew for p in ps) > 100
    qs = tuning.sample_params(56, 400, **tuning.WORKLOAD_RANGES["batched"])
    assert all(1e3 <= q.nr_rows <= 2e5 and q.skew <= 100 for q in qs)


def test_load_corpora_stacks_and_checks_variants(tmp_path):
    from spmvdare.train_selector import load_corpora
    rng = np.random.default_rng(0)
    paths = []
    for k, rows in enumerate((3, 5)):
        p = tmp_path / ("c%d.npz" % k)
        np.savez_compressed(p, features=rng.random((rows, 14)), times=rng.random((rows, len(NAMES))) + 0.1,
                            variants=np.array(NAMES))
        paths.append(str(p))
    X, T = load_corpora(paths)
    assert X.shape == (8, 14) and T.shape == (8, len(NAMES))
    np.testing.assert_array_equal(T[3:], np.load(paths[1])["times"])
    bad = tmp_path / "bad.npz"
    np.savez_compressed(bad, features=np.zeros((1, 14)), times=np.ones((1, len(NAMES))),
                        variants=np.array(list(reversed(NAMES))))
    with pytest.raises(SystemExit):
        load_corpora([paths[0], str(bad)])


def test_consistency_spread():
    """The stream's fallback trigger: timings whose ratios are the predicted
    ones give spread 0 whatever the absolute time; one timing twice as slow
    as predicted relative to the others gives log 2."""
    from spmvdare.stream import consistency_spread
    pred = np.array([0.0, 0.1, 0.4])
    t = 3.0 * np.exp(pred)
    assert abs(consistency_spread(pred, t)) < 1e-12
    assert abs(consistency_spread(pred, t * 7.5)) < 1e-12
    t2 = t.copy()
    t2[2] *= 2.0
    assert abs(consistency_spread(pred, t2) - np.log(2.0)) < 1e-12
    t3 = t.copy()
    t3[0] *= 1.25  # the top prediction slower than predicted
    assert abs(consistency_spread(pred, t3) - np.log(1.25)) < 1e-12


def test_batch_choice_div_margin():
    """Inside a batched launch a DIV-RF choice must beat the best other
    allowed variant by the margin; otherwise that variant runs (with the
    batch's CSR block size)."""
    allowed = tuning.batch_allowed()
    div = tuning.VARIANTS.index(("div_rf", {"div_packet_vals": 1024, "div_threads": 1, "unroll": 8,
                                            "xcd_remap": 1}))
    csr128 = tuning.VARIANTS.index(("csr_vec", {"block_nnz": 128, "lanes_per_row": 0}))
    csr256 = tuning.VARIANTS.index(("csr_vec", {"block_nnz": 256, "lanes_per_row": 0}))
    csr512 = tuning.VARIANTS.index(("csr_vec", {"block_nnz": 512, "lanes_per_row": 0}))
    pred = np.full(len(NAMES), 2.0)
    pred[div], pred[csr128] = 0.0, np.log(1.05)        # DIV predicted 5 % faster: not enough
    assert tuning.batch_choice(pred, allowed, 256, 0.1) == csr256
    pred[csr128] = np.log(1.2)                           # 20 % faster: DIV kept
    assert tuning.batch_choice(pred, allowed, 256, 0.1) == div
    assert tuning.batch_choice(pred, allowed, 256, 0.5) == csr256
    pred[:] = 2.0
    pred[csr512], pred[csr256] = 0.0, 0.1                # block 512 is not allowed in a batch
    assert tuning.batch_choice(pred, allowed, 256, 0.1) == csr256


def test_batch_variant_mapping():
    """Inside a batch every CSR choice runs as the default CSR unit (paired
    loads, automatic reduction width) with blocks of at least the batch's
    block size; min_block_nnz 0 keeps the block size; other formats are
    kept."""
    V = tuning.VARIANTS
    idx = lambda k, kw: V.index((k, kw))  # noqa: E731
    d256 = idx("csr_vec", {"block_nnz": 256, "lanes_per_row": 0})
    for kw in ({"block_nnz": 64, "lanes_per_row": 1}, {"block_nnz": 128, "lanes_per_row": 4},
               {"block_nnz": 256, "lanes_per_row": 0, "narrow_loads": 1},
               {"block_nnz": 256, "lanes_per_row": 4, "xcd_remap": 1}, {"block_nnz": 256, "lanes_per_row": 0}):
        assert tuning.batch_variant(idx("csr_vec", kw), 256) == d256
    assert tuning.batch_variant(idx("csr_vec", {"block_nnz": 512, "lanes_per_row": 1}), 256) == \
        idx("csr_vec", {"block_nnz": 512, "lanes_per_row": 0})
    assert tuning.batch_variant(idx("csr_vec", {"block_nnz": 64, "lanes_per_row": 1}), 0) == \
        idx("csr_vec", {"block_nnz": 64, "lanes_per_row": 0})
    for k, kw in V:
        if k != "csr_vec":
            assert tuning.batch_variant(V.index((k, kw)), 256) == V.index((k, kw))
