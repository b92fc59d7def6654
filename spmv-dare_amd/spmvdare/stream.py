"""End-to-end streaming of generated matrices through the selector and the
batched launch.

A stream of generator parameter sets is processed in chunks that fit HBM:
the host generates chunk i + 1 (a worker thread; the generator releases the
interpreter lock) while the GPU takes chunk i: each matrix is uploaded, its
structure statistics are computed on the device, the learned selector
predicts every variant's slowdown, and the matrix is converted to the chosen
variant: the `refine` best predicted variants are timed and the fastest
is kept, and when those timings contradict the predictions (their ratios
differ from the predicted ones by more than `fallback_spread`: the selector
is off on this matrix) the choice falls back to exhaustive autotuning over
every variant the batched launch can run. The chunk then
runs as one persistent batched launch replayed from a hipGraph, and is
compared on the same GPU with plain per-matrix CSR-vector launches. Every
member's y is checked against the device csr_naive of its matrix (a
sequential left-to-right sum per row, bit-identical to the reference's
csr_naive semantics) at 1e-13 normwise. The first `tuned` matrices are also
tuned exhaustively, which judges the selector's regret.

Nothing here has a counterpart in the reference, whose driver runs one
matrix per process (run.sh:116-151).
"""
import math
import threading
import time

import numpy as np

from . import gen, hip, tuning


def alg_bytes(m, n, nnz):
    return 12 * nnz + 4 * (m + 1) + 8 * n + 8 * m


def _rel(y, yr):
    nrm = np.linalg.norm(yr)
    return float(np.linalg.norm(y - yr) / nrm) if nrm > 0 else float(np.linalg.norm(y))


def consistency_spread(pred, times):
    """How far timings contradict the selector's predictions. The selector
    predicts log(t_v / t_best) for every variant v, so each timed variant
    gives an estimate log(t_v) - pred_v of log t_best; when the predictions
    are right these agree. Returns the range of those estimates (natural
    log): 0 when the timed ratios are exactly the predicted ones."""
    e = np.log(np.asarray(times, dtype=np.float64)) - np.asarray(pred, dtype=np.float64)
    return float(e.max() - e.min())


def choose(sel, dm, p, allowed, fallback_spread, tune, refine=3, warmup=1, iters=5):
    """The variant a matrix runs as, and what was learned choosing it.

    The `refine` best predictions are timed and the fastest is kept. When
    their timings disagree with the predictions by more than fallback_spread
    (consistency_spread above log1p(fallback_spread): the selector is off on
    this matrix, e.g. outside what it was trained on), the choice falls back
    to exhaustive autotuning over `allowed`. With refine = 1 the top
    prediction is kept untimed and nothing falls back. With `tune` every
    allowed variant is timed first (to judge the regret) and the policy reads
    its timings from that run. Returns (variant, info): info["fallback"],
    info["spread"], info["top1"], and info["times"] (ms over `allowed`, NaN
    elsewhere) when exhaustive timings were taken."""
    feats = tuning.features(p, dm.stats())
    pred = sel.predict(feats)[0].astype(np.float64)
    pred[~allowed] = np.inf
    order = np.argsort(pred)
    info = {"top1": int(order[0]), "fallback": False, "spread": 0.0}
    t = None
    idx = np.flatnonzero(allowed)
    if tune:
        t = np.full(len(tuning.VARIANTS), np.nan)
        t[idx] = tuning.autotune(dm, variants=idx, warmup=warmup, iters=iters)
        info["times"] = t
    top = [int(i) for i in order[:max(1, refine)] if np.isfinite(pred[i])]
    if len(top) == 1:
        return top[0], info
    tt = t[top] if t is not None else tuning.autotune(dm, variants=top, warmup=warmup, iters=iters)
    info["spread"] = consistency_spread(pred[top], tt)
    if info["spread"] > math.log1p(fallback_spread):
        info["fallback"] = True
        if t is None:
            t = np.full(len(tuning.VARIANTS), np.nan)
            t[idx] = tuning.autotune(dm, variants=idx, warmup=warmup, iters=iters)
            info["times"] = t
        return int(np.nanargmin(t)), info
    return top[int(np.argmin(tt))], info


def run_stream(params, count, chunk, selector, tuned=0, fallback_spread=0.3, steps=10, batch_bnnz=256,
               queues=256, tile=0, log=None, keep_y=False, solo_nnz=0, refine=3):
    """Process the first `count` generatable matrices of `params` (a list of
    gen.GenParams) as described in the module docstring. Returns a dict of
    totals; with keep_y, also the list of (params, batched y) for checks."""
    allowed = tuning.batch_allowed()
    csr = tuning.DEFAULT_CSR
    stream = iter(params)

    def produce(n):
        out = []
        for p in stream:
            try:
                out.append((p, gen.generate(p, stats=False)))
            except gen.GenError:
                continue
            if len(out) == n:
                break
        return out

    t_start = time.perf_counter()
    done = 0
    tot = {"nnz": 0, "bytes": 0, "ms_batched": 0.0, "ms_plain": 0.0}
    worst, worst_plain, chosen, fallbacks, regrets, kept = 0.0, 0.0, [], 0, [], []
    pending = produce(min(chunk, count))
    while pending:
        box = {}
        want = min(chunk, count - done - len(pending))
        worker = None
        if want > 0:
            def fill(n=want, box=box):
                try:
                    box["m"] = produce(n)
                except BaseException as e:  # handed to the main thread below
                    box["error"] = e

            worker = threading.Thread(target=fill)
            worker.start()
        mats = []
        for p, a in pending:
            dm = tuning.DeviceCSR(a, p.seed + 1)
            v, info = choose(selector, dm, p, allowed, fallback_spread, done + len(mats) < tuned, refine=refine)
            fallbacks += info["fallback"]
            if "times" in info and done + len(mats) < tuned:
                t = info["times"]
                best = np.nanmin(t)
                regrets.append((float(t[info["top1"]] / best - 1.0), float(t[v] / best - 1.0)))
            chosen.append(tuning.variant_name(v))
            f_sel = dm.convert(tuning.batch_variant(v, batch_bnnz))
            f_csr = dm.convert(csr)
            mats.append((p, a, dm, f_sel, f_csr, hip.DeviceArray(max(a.m, 1))))
            tot["nnz"] += a.nnz
            tot["bytes"] += alg_bytes(a.m, a.n, a.nnz)
            dm.a = None
        batch, base = hip.Batch(tile), hip.Batch(tile)
        batch.set_queues(queues)
        if solo_nnz:
            batch.set_solo_nnz(solo_nnz)
        for p, a, dm, f_sel, f_csr, dy_csr in mats:
            batch.add(f_sel, dm.dx, dm.dy)
            base.add(f_csr, dm.dx, dy_csr)
        batch.finalize(use_graph=True)
        base.finalize(use_graph=False)
        tot["ms_batched"] += float(np.median(batch.time(2, steps)))
        tot["ms_plain"] += float(np.median(base.time(1, steps, per_matrix=True)))
        batch.run()
        batch.synchronize()
        for p, a, dm, f_sel, f_csr, dy_csr in mats:
            yb, yc = dm.dy.download()[:a.m], dy_csr.download()[:a.m]
            f_ref = hip.csr_to_format_device(dm.rp, dm.col, dm.val, dm.m, dm.n, dm.nnz, kind="csr_naive")
            f_ref.spmv_device(dm.dx, dy_csr)
            hip.synchronize()
            yr = dy_csr.download()[:a.m]
            f_ref.destroy()
            worst = max(worst, _rel(yb, yr))
            worst_plain = max(worst_plain, _rel(yc, yr))
            if keep_y:
                kept.append((p, yb))
        done += len(mats)
        batch.destroy()
        base.destroy()
        for m_ in mats:
            m_[3].destroy()
            m_[4].destroy()
        del mats
        if log:
            log("stream: %d/%d matrices, %.0f s, batched %.1f ms vs per-matrix CSR %.1f ms so far, %d fallbacks" % (
                done, count, time.perf_counter() - t_start, tot["ms_batched"], tot["ms_plain"], fallbacks))
        if worker is not None:
            worker.join()
        if "error" in box:
            raise RuntimeError("stream: host generation of the next chunk failed") from box["error"]
        pending = box.get("m", [])
    t_b, t_p = tot["ms_batched"], tot["ms_plain"]
    r = np.array(regrets) if regrets else np.zeros((0, 2))
    out = {
        "matrices": done, "nnz": tot["nnz"], "alg_bytes": tot["bytes"], "chunk": chunk,
        "batched_ms": t_b, "plain_per_matrix_csr_ms": t_p,
        "matrices_per_s": done / (t_b * 1e-3) if t_b else None,
        "plain_matrices_per_s": done / (t_p * 1e-3) if t_p else None,
        "batched_gbps": tot["bytes"] / (t_b * 1e-3) / 1e9 if t_b else None,
        "speedup_vs_plain_csr": t_p / t_b if t_b else None,
        "max_normwise_err_vs_csr_naive": worst,
        "max_normwise_err_vs_csr_naive_plain_csr": worst_plain,
        "fallback_spread": fallback_spread, "fallbacks": fallbacks, "refine": refine,
        "regret": {"matrices": len(regrets),
                   "top1_mean": float(r[:, 0].mean()) if len(r) else None,
                   "top1_max": float(r[:, 0].max()) if len(r) else None,
                   "policy_mean": float(r[:, 1].mean()) if len(r) else None,
                   "policy_max": float(r[:, 1].max()) if len(r) else None,
                   "policy_frac_le_5pct": float((r[:, 1] <= 0.05).mean()) if len(r) else None},
        "chosen_variants": {k: chosen.count(k) for k in sorted(set(chosen))},
        "wall_s": time.perf_counter() - t_start,
    }
    if keep_y:
        out["y"] = kept
    return out
