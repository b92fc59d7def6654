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
    for f in ("corpus_train.npz", "corpus_test.npz"):
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
