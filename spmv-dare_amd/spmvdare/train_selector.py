"""Train the format selector on corpora and judge it on held-out corpora.

    python -m spmvdare.train_selector --train train.npz [more.npz ...] \
        --test test.npz [more_test.npz ...] --out selector.json

The training corpora are concatenated. Reports, on each held-out corpus
separately, the regret t_chosen / t_best - 1 against exhaustive autotuning
for: the MLP's top choice, the MLP's top-3 refined by timing them (the
fallback), the single variant best on average over the training set (the
method without the learned part), and the default binned-CSR variant.
"""
import argparse
import json
import os
import sys

import numpy as np

from . import hip, tuning


def load_corpora(paths):
    """Features and timings of several corpora (same variant list), stacked."""
    names = [tuning.variant_name(i) for i in range(len(tuning.VARIANTS))]
    X, T = [], []
    for path in paths:
        c = np.load(path)
        if list(c["variants"]) != names:
            raise SystemExit("%s was timed on another variant list" % path)
        X.append(c["features"])
        T.append(c["times"])
    return np.concatenate(X), np.concatenate(T)


def regrets(times, choice):
    best = times.min(axis=1)
    return times[np.arange(len(times)), choice] / best - 1.0


def summary(r):
    r = np.asarray(r)
    return {"mean": float(r.mean()), "median": float(np.median(r)), "p95": float(np.percentile(r, 95)),
            "max": float(r.max()), "frac_le_5pct": float((r <= 0.05).mean())}


def evaluate(sel, Xte, Tte, Ttr, k=3):
    pred = sel.predict(Xte)
    order = np.argsort(pred, axis=1)
    top1 = order[:, 0]
    topk = order[:, :k]
    refined = topk[np.arange(len(topk)), np.argmin(Tte[np.arange(len(topk))[:, None], topk], axis=1)]
    static = int(np.argmin(np.log(Ttr / Ttr.min(axis=1, keepdims=True)).mean(axis=0)))
    default = tuning.VARIANTS.index(("csr_vec", {"block_nnz": 256, "lanes_per_row": 0}))
    return {
        "mlp_top1": summary(regrets(Tte, top1)),
        "mlp_top%d_refined" % k: summary(regrets(Tte, refined)),
        "static_best_variant": dict(summary(regrets(Tte, np.full(len(Tte), static))),
                                    variant=tuning.variant_name(static)),
        "default_csr_vec": summary(regrets(Tte, np.full(len(Tte), default))),
        "oracle_exhaustive": summary(np.zeros(len(Tte))),
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--train", nargs="+", required=True)
    ap.add_argument("--test", nargs="+", required=True)
    ap.add_argument("--out", required=True)
    ap.add_argument("--epochs", type=int, default=400)
    ap.add_argument("--lr", type=float, default=2e-3)
    args = ap.parse_args()
    hip.init(0)
    X, T = load_corpora(args.train)
    sel = tuning.Selector()
    losses = sel.fit(X, T, epochs=args.epochs, lr=args.lr)
    print("train loss: first %.4f last %.4f" % (losses[0], losses[-1]), file=sys.stderr, flush=True)
    res = {"train_matrices": int(len(T)), "final_train_loss": float(losses[-1]), "test": {}}
    for path in args.test:
        te = np.load(path)
        r = evaluate(sel, te["features"], te["times"], T)
        r["test_matrices"] = int(len(te["times"]))
        res["test"][os.path.basename(path)] = r
    sel.save(args.out)
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
