"""Build a selector corpus: generated matrices x the tuning variants, each
timed on the GPU by the exhaustive autotuner, with the 14 selector features.

    python -m spmvdare.corpus --count 4096 --seed 33 --out corpus.npz

When variants are added, an existing corpus gains their columns without
timing the others again: time only the new variants on the same draws, then
merge (columns in variant-list order):

    python -m spmvdare.corpus --seed 33 --draws-from corpus.npz --variants 40,41 --out new.npz
    python -m spmvdare.corpus --merge corpus.npz new.npz --out corpus2.npz

Matrices are sampled from the batched workload's parameter ranges (rows = cols
log-uniform in [1e3, 2e5], avg 4..64, std 0.25..2 x avg, normal / gamma,
random / diagonal, bw 0.01..1, skew 1..100, neighbours 0..8, cross-row
similarity 0..0.9), or with `--ranges stream` from the streaming workload's
(rows log-uniform in [1e4, 4e6], avg 2..128, gamma rows, skew 1..1000, the
rest as above):

    python -m spmvdare.corpus --ranges stream --seed 56 --budget 900 --out stream_train.npz

A parameter set the generator rejects (a cross-row copy collision, where the
reference exits) is replaced by the next draw.
"""
import argparse
import sys
import time

import numpy as np

from . import gen, hip, tuning


def build(count, seed, warmup=2, iters=8, log=None, ranges=None, budget_s=None, start=0, stop=None,
          draws=None, variants=None, checkpoint=None):
    """Time `count` matrices drawn from sample_params(seed), beginning at draw
    `start`, or (stop given) every matrix of the draws [start, stop), so that
    a corpus can be built in several sessions and concatenated; or (draws
    given) exactly those draws. `variants`: indices into tuning.VARIANTS to
    time (default all). `checkpoint(arrays)` is called with the corpus so far
    every 32 matrices; `log` gets a progress line every 8."""
    if draws is not None:
        draws = [int(d) for d in draws]
        params = tuning.sample_params(seed, max(draws) + 1 if draws else 0, **(ranges or {}))
        order = draws
        count = len(draws)
    else:
        params = tuning.sample_params(seed, stop if stop is not None else start + count * 2, **(ranges or {}))
        if stop is not None:
            count = stop - start
        order = range(start, len(params))
    feats, times, lines, nnzs, pidx = [], [], [], [], []
    t0 = time.time()
    k = 0
    for i in order:
        p = params[i]
        if len(lines) == count:
            break
        if budget_s is not None and time.time() - t0 > budget_s:
            if log:
                log("corpus: time budget reached after %d matrices" % k)
            break
        try:
            a = gen.generate(p, stats=False)
        except gen.GenError:
            if draws is not None:
                raise RuntimeError("draw %d no longer generates" % i)
            continue
        dm = tuning.DeviceCSR(a, p.seed + 1)
        st = dm.stats()
        feats.append(tuning.features(p, st))
        times.append(tuning.autotune(dm, variants=variants, warmup=warmup, iters=iters))
        lines.append(p.line())
        nnzs.append(a.nnz)
        pidx.append(i)
        del dm
        k += 1
        if log and (k % 8 == 0 or k == count):
            log("corpus %d/%d matrices, %.0f s" % (k, count, time.time() - t0))
        if k % 32 == 0 or k == count:
            if checkpoint:
                checkpoint((np.array(feats), np.array(times), np.array(lines), np.array(nnzs), np.array(pidx)))
    return np.array(feats), np.array(times), np.array(lines), np.array(nnzs), np.array(pidx)


def merge(base, add, out):
    """Columns of `add` (timed for the variants it names) joined to `base`,
    row by row on the same draws, in tuning.VARIANTS order."""
    a, b = np.load(base), np.load(add)
    if list(a["draw"]) != list(b["draw"]) or list(a["lines"]) != list(b["lines"]):
        raise SystemExit("merge: the corpora hold different matrices")
    names = [tuning.variant_name(i) for i in range(len(tuning.VARIANTS))]
    cols = {n: a["times"][:, j] for j, n in enumerate(a["variants"])}
    cols.update({n: b["times"][:, j] for j, n in enumerate(b["variants"])})
    missing = [n for n in names if n not in cols]
    if missing:
        raise SystemExit("merge: no timings for %s" % ", ".join(missing))
    np.savez_compressed(out, features=a["features"], times=np.stack([cols[n] for n in names], axis=1),
                        lines=a["lines"], nnz=a["nnz"], draw=a["draw"], variants=np.array(names))


def main():
    if len(sys.argv) > 1 and sys.argv[1] == "--merge":
        mp = argparse.ArgumentParser()
        mp.add_argument("--merge", nargs=2, required=True)
        mp.add_argument("--out", required=True)
        margs = mp.parse_args()
        merge(margs.merge[0], margs.merge[1], margs.out)
        return
    ap = argparse.ArgumentParser()
    ap.add_argument("--count", type=int, default=4096)
    ap.add_argument("--seed", type=int, default=33)
    ap.add_argument("--out", required=True)
    ap.add_argument("--iters", type=int, default=8)
    ap.add_argument("--budget", type=float, default=None, help="stop after this many seconds")
    ap.add_argument("--start", type=int, default=0, help="first parameter draw to use")
    ap.add_argument("--stop", type=int, default=None,
                    help="time every matrix of the draws [start, stop) (instead of --count matrices)")
    ap.add_argument("--draws-from", default=None, help="time the draws of this corpus (same --seed)")
    ap.add_argument("--variants", default=None, help="comma-separated variant indices (default all)")
    ap.add_argument("--ranges", default="batched", choices=sorted(tuning.WORKLOAD_RANGES),
                    help="parameter ranges of the draws: the batched or the streaming workload's")
    args = ap.parse_args()
    draws = np.load(args.draws_from)["draw"] if args.draws_from else None
    vidx = [int(v) for v in args.variants.split(",")] if args.variants else None
    hip.init(0)

    def log(s):
        print(s, file=sys.stderr, flush=True)

    names = [tuning.variant_name(i) for i in (vidx if vidx is not None else range(len(tuning.VARIANTS)))]

    def save(arrays):
        X, T, L, N, I = arrays
        np.savez_compressed(args.out, features=X, times=T, lines=L, nnz=N, draw=I, variants=np.array(names),
                            ranges=np.array(args.ranges), seed=np.array(args.seed))

    arrays = build(args.count, args.seed, iters=args.iters, log=log, budget_s=args.budget,
                   start=args.start, stop=args.stop, draws=draws, variants=vidx,
                   ranges=tuning.WORKLOAD_RANGES[args.ranges], checkpoint=save)
    save(arrays)
    log("wrote %s: %d matrices x %d variants" % (args.out, arrays[1].shape[0], len(names)))


if __name__ == "__main__":
    main()
