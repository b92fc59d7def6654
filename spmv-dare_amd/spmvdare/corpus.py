"""Build a selector corpus: generated matrices x the 40 tuning variants, each
timed on the GPU by the exhaustive autotuner, with the 14 selector features.

    python -m spmvdare.corpus --count 4096 --seed 33 --out corpus.npz

Matrices are sampled from the batched workload's parameter ranges (rows = cols
log-uniform in [1e3, 2e5], avg 4..64, std 0.25..2 x avg, normal / gamma,
random / diagonal, bw 0.01..1, skew 1..100, neighbours 0..8, cross-row
similarity 0..0.9). A parameter set the generator rejects (a cross-row copy
collision, where the reference exits) is replaced by the next draw.
"""
import argparse
import sys
import time

import numpy as np

from . import gen, hip, tuning


def build(count, seed, warmup=2, iters=8, log=None, ranges=None, budget_s=None, start=0, stop=None):
    """Time `count` matrices drawn from sample_params(seed), beginning at draw
    `start`, or (stop given) every matrix of the draws [start, stop), so that
    a corpus can be built in several sessions and concatenated."""
    params = tuning.sample_params(seed, stop if stop is not None else start + count * 2, **(ranges or {}))
    if stop is not None:
        count = stop - start
    feats, times, lines, nnzs, pidx = [], [], [], [], []
    t0 = time.time()
    k = 0
    for i in range(start, len(params)):
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
            continue
        dm = tuning.DeviceCSR(a, p.seed + 1)
        st = dm.stats()
        feats.append(tuning.features(p, st))
        times.append(tuning.autotune(dm, warmup=warmup, iters=iters))
        lines.append(p.line())
        nnzs.append(a.nnz)
        pidx.append(i)
        del dm
        k += 1
        if log and (k % 32 == 0 or k == count):
            log("corpus %d/%d matrices, %.0f s" % (k, count, time.time() - t0))
    return np.array(feats), np.array(times), np.array(lines), np.array(nnzs), np.array(pidx)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--count", type=int, default=4096)
    ap.add_argument("--seed", type=int, default=33)
    ap.add_argument("--out", required=True)
    ap.add_argument("--iters", type=int, default=8)
    ap.add_argument("--budget", type=float, default=None, help="stop after this many seconds")
    ap.add_argument("--start", type=int, default=0, help="first parameter draw to use")
    ap.add_argument("--stop", type=int, default=None,
                    help="time every matrix of the draws [start, stop) (instead of --count matrices)")
    args = ap.parse_args()
    hip.init(0)

    def log(s):
        print(s, file=sys.stderr, flush=True)

    X, T, L, N, I = build(args.count, args.seed, iters=args.iters, log=log, budget_s=args.budget,
                          start=args.start, stop=args.stop)
    np.savez_compressed(args.out, features=X, times=T, lines=L, nnz=N, draw=I,
                        variants=np.array([tuning.variant_name(i) for i in range(len(tuning.VARIANTS))]))
    log("wrote %s: %d matrices x %d variants" % (args.out, T.shape[0], T.shape[1]))


if __name__ == "__main__":
    main()
