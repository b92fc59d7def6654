#!/usr/bin/env python3
"""Benchmark: FP64 SpMV on one MI355X, algorithmic GB/s and % of HBM peak.

Default workload: the 8,388,608 x 8,388,608 generated matrix (avg 32 nnz/row,
gamma, std 48, skew 20, diagonal placement, bw 0.05, 4 neighbours,
cross-row similarity 0.5, seed 7; 268,448,832 nnz). Every drop-in format is
converted on the device and timed; the line reports the best one. A "step"
is one y = A x over the whole matrix with x, y and A resident in HBM.

    python bench.py [--gpus 1] [--steps 128] [--warmup 10]

One JSON line on stdout; progress on stderr.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "spmv-dare_amd"))

HBM_PEAK_GBS = 8000.0
WORKLOADS = {
    "spmv8m": ("8388608 8388608 32 48 gamma diagonal 0.05 20 4 0.5 7",
               "8,388,608^2 generated matrix: avg 32 nnz/row, gamma std 48, skew 20, "
               "diagonal, bw 0.05, 4 neighbours, cross-row 0.5, seed 7"),
    "spmv100k": ("100000 100000 16 4 normal random 0.3 1.5 1 0.2 14",
                 "100,000^2 generated matrix: avg 16 nnz/row, normal std 4, skew 1.5, "
                 "random, bw 0.3, 1 neighbour, cross-row 0.2, seed 14"),
}
CANDIDATES = [
    ("csr_vec", {"block_nnz": 256}),
    ("csr_vec", {"block_nnz": 256, "xcd_remap": 1}),
    ("sell_c_s", {"sigma": 256, "lanes_per_row": 4, "unroll": 4}),
    ("sell_c_s", {"sigma": 256}),
    ("sell_c_s", {"sigma": 4096, "lanes_per_row": 4, "unroll": 4}),
    ("sell_c_s", {"sigma": 1}),
    ("sell", {"lanes_per_row": 4, "unroll": 4}),
    ("div_rf", {"div_packet_vals": 1024, "div_threads": 64, "unroll": 8, "xcd_remap": 1}),
    ("csr_naive", {}),
]


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def alg_bytes(m, n, nnz):
    """12 nnz + 4 (m + 1) + 8 n + 8 m: FP64 value + int32 column per nnz,
    int32 row pointer, x read once, y written once."""
    return 12 * nnz + 4 * (m + 1) + 8 * n + 8 * m


REF_CPU_FORMATS = ("csr_naive", "csr_vec", "sell_c_s", "div_rf")


def ref_cpu_exe(fmt):
    return os.path.join(ROOT, "oracle", "_ref", "cpu_%s_d.exe" % fmt)


def host_cpus():
    """One logical CPU per physical core among the CPUs this process may run
    on (sysfs topology), and the CPU model: the 1:1 pinning of the
    reference's calc_cpu_pinning (config.sh:16-31) with one thread per core."""
    allowed = sorted(os.sched_getaffinity(0))
    first = {}
    for c in allowed:
        base = "/sys/devices/system/cpu/cpu%d/topology/" % c
        try:
            key = (open(base + "physical_package_id").read().strip(), open(base + "core_id").read().strip())
        except OSError:
            key = ("?", str(c))
        first.setdefault(key, c)
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.lower().startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return sorted(first.values()), len(allowed), model


def cpu_baseline_reference(line, m, n, nnz, timeout_s):
    """The reference's own CPU formats (csr_naive, csr_vec, div_rf; sell_c_s
    with its own conversion, analyzer and scheduler around a portable kernel
    in place of its RISC-V one), each an executable built from its sources
    with its own driver (src/bench.cpp + src/bench_spmv.cpp). The driver
    generates the same matrix from the same parameter line, then times >= 64
    iterations and >= 2 s and prints the median (bench_spmv.cpp:313-359).
    One OpenMP thread per physical core, pinned with GOMP_CPU_AFFINITY as
    run.sh:68-72 does."""
    import re
    import subprocess
    cpus, logical, model = host_cpus()
    env = dict(os.environ, USE_ARTIFICIAL_MATRICES="1", USE_PROCESSES="0", GPU_KERNEL="0",
               USE_RCM_REORDERING="0", COOLDOWN="0", NUM_PROCESSES="1", OMP_NUM_THREADS=str(len(cpus)),
               GOMP_CPU_AFFINITY=",".join(str(c) for c in cpus), OMP_DYNAMIC="false",
               CSRCV_NUM_PACKET_VALS="16384", DIV_VC_TOLERANCE="0", DIV_PACKET_REORDERING="0",
               DIV_STEALING_ENABLED="1")
    env.pop("OMP_PROC_BIND", None)
    env.pop("OMP_PLACES", None)
    B = alg_bytes(m, n, nnz)
    per = {}
    for fmt in REF_CPU_FORMATS:
        exe = ref_cpu_exe(fmt)
        if not os.path.exists(exe):
            continue
        env["PROGG"] = fmt
        t0 = time.perf_counter()
        try:
            r = subprocess.run([exe] + line.split(), env=env, capture_output=True, text=True, timeout=timeout_s)
        except subprocess.TimeoutExpired:
            log("cpu baseline %s: no result within %d s" % (fmt, timeout_s))
            continue
        mt = re.search(r"time iter: min=([0-9.eE+-]+), median=([0-9.eE+-]+)", r.stdout)
        ml = re.search(r"number of loops = (\d+)", r.stdout)
        if r.returncode != 0 or not mt:
            log("cpu baseline %s: driver exit %d" % (fmt, r.returncode))
            continue
        med = float(mt.group(2))
        per[fmt] = {"gbps": B / med / 1e9, "gflops": 2 * nnz / med / 1e9, "median_ms": med * 1e3,
                    "min_ms": float(mt.group(1)) * 1e3, "loops": int(ml.group(1)) if ml else None,
                    "wall_s": time.perf_counter() - t0,
                    "kernel": "reference" if fmt != "sell_c_s" else "reference conversion and scheduler, "
                                                                    "portable lane-loop kernel"}
        log("cpu baseline %-9s %7.1f GB/s  median %.2f ms  (%d loops, %.0f s with generation)" % (
            fmt, per[fmt]["gbps"], med * 1e3, per[fmt]["loops"] or 0, per[fmt]["wall_s"]))
    if not per:
        return None
    best = max(per, key=lambda k: per[k]["gbps"])
    return {"value": round(per[best]["gbps"], 2), "unit": "GB/s", "cores": len(cpus), "kind": "reference",
            "sample": "reference CPU %s (best of %s; the reference's own driver, >= 64 iterations and >= 2 s, "
                      "median) on the full matrix, one OpenMP thread per physical core pinned 1:1 "
                      "(%d cores of %d logical CPUs, %s), AVX2 build" % (
                          best, ", ".join(sorted(per)), len(cpus), logical, model),
            "format": best, "gflops": round(per[best]["gflops"], 2), "cpu_model": model,
            "physical_cores": len(cpus), "logical_cpus": logical,
            "formats": {k: {kk: (round(vv, 3) if isinstance(vv, float) else vv) for kk, vv in v.items()}
                        for k, v in per.items()}}


def cpu_baseline_port(a, x, budget_s):
    """The oracle's OpenMP csr_naive (equal-row split as csr.cpp:74) on the
    host cores, on the same matrix: one warm-up call, then iterations until
    the time budget (at least 3), median time. Used where the reference's
    CPU executables were not built."""
    import numpy as np
    import pyoracle
    cpus, logical, model = host_cpus()
    y, t = pyoracle.csr_naive_omp(a.row_ptr, a.col_ind, a.values, x, threads=len(cpus))
    times = []
    t_end = time.perf_counter() + budget_s
    while len(times) < 3 or (time.perf_counter() < t_end and len(times) < 128):
        t0 = time.perf_counter()
        pyoracle.csr_naive_omp(a.row_ptr, a.col_ind, a.values, x, threads=len(cpus), y=y)
        times.append(time.perf_counter() - t0)
    med = float(np.median(times))
    return {"value": round(alg_bytes(a.m, a.n, a.nnz) / med / 1e9, 2), "unit": "GB/s", "cores": t,
            "kind": "port",
            "sample": "oracle csr_naive (OpenMP, equal-row split) on the full matrix, "
                      "%d timed iterations after 1 warm-up, median %.2f ms (%s)" % (len(times), med * 1e3, model),
            "gflops": 2 * a.nnz / med / 1e9, "cpu_model": model}


def check_leg(a, x, ys, best_key):
    """The oracle as the checker: y of the reference's csr_naive semantics
    (sequential left-to-right row sums; OpenMP over rows, which does not change
    any row's sum) and the double-double oracle, against every GPU y.
    Returns {format key: normwise relative error vs csr_naive} and the best
    format's error against the double-double y."""
    import numpy as np
    import pyoracle
    cpus, _, _ = host_cpus()
    t0 = time.perf_counter()
    y_ref, _ = pyoracle.csr_naive_omp(a.row_ptr, a.col_ind, a.values, x, threads=len(cpus))
    nrm = float(np.linalg.norm(y_ref))
    errs = {k: float(np.linalg.norm(y - y_ref) / nrm) for k, y in ys.items()}
    hi, lo = pyoracle.csr_dd(a.row_ptr, a.col_ind, a.values, x)
    dd = float(np.linalg.norm((ys[best_key] - hi) - lo) / np.linalg.norm(hi))
    dd_naive = float(np.linalg.norm((y_ref - hi) - lo) / np.linalg.norm(hi))
    log("check: %d formats against the oracle in %.1f s; worst normwise error %.2e; best format vs "
        "double-double %.2e (csr_naive itself %.2e)" % (len(errs), time.perf_counter() - t0, max(errs.values()),
                                                         dd, dd_naive))
    return errs, dd, dd_naive


PROFILED = {  # the variant each profile summary's byte counters were taken on
    ("spmv8m", "csr_vec"): {"block_nnz": 256, "xcd_remap": 1},
    ("spmv8m", "sell_c_s"): {"sigma": 256, "lanes_per_row": 4, "unroll": 4},
    ("spmv8m", "div_rf"): {"div_packet_vals": 1024, "div_threads": 64, "unroll": 8, "xcd_remap": 1},
}


def profiled_traffic(workload, fmt, params):
    """HBM bytes per launch of this format on this workload from the committed
    rocprofv3 summary (FETCH_SIZE and WRITE_SIZE, calibrated), or None when
    the summary does not cover exactly this variant."""
    import glob
    import re
    if PROFILED.get((workload, fmt)) != params:
        return None
    for path in sorted(glob.glob(os.path.join(ROOT, "profiles", "%s_*.md" % workload))):
        for line in open(path):
            m = re.match(r"^\| %s \| [^|]+ \| \d+ \| \d+ \| (\d+) \| [\d.]+ \|\s*$" % re.escape(fmt), line)
            if m:
                return int(m.group(1))
    return None


def parse_format(tok):
    """kind[:sigma] or kind:key=value;key=value"""
    k, _, rest = tok.partition(":")
    if not rest:
        return k, {}
    if "=" not in rest:
        return k, {"sigma": int(rest)}
    return k, {kv.split("=")[0]: int(kv.split("=")[1]) for kv in rest.split(";")}


def probe(hip, np):
    """Stream-read and copy bandwidth of 2 GiB buffers: the measured ceiling
    and the known traffic for calibrating the profiler's byte counters."""
    n = 1 << 28  # doubles = 2 GiB
    src = hip.DeviceArray(n)
    dst = hip.DeviceArray(n)
    src.fill_bytes(0)
    out = {}
    for mode, name, nbytes in ((0, "stream_read", 8 * n), (1, "stream_copy", 16 * n), (2, "stream_read_lds", 8 * n)):
        ms = hip.bandwidth_probe(mode, src, dst if mode == 1 else None, 20)
        out[name] = {"bytes": nbytes, "median_ms": float(np.median(ms)),
                     "gbps": nbytes / (float(np.median(ms)) * 1e-3) / 1e9}
        log("%s: %.3f ms, %.1f GB/s" % (name, out[name]["median_ms"], out[name]["gbps"]))
    del src, dst
    # the cost of one 64-lane load by pattern, on a buffer that stays in the
    # L1 (16 KiB) and one that stays in the L2 (2 MiB): cycles per CU at 2.4 GHz
    names = ["contiguous 4 B", "contiguous 8 B", "contiguous 16 B", "8 B gather, 64 lines",
             "8 B gather, 64 lines, half the lanes", "16 B gather, 64 lines", "8 B gather, 16 lines",
             "8 B gather, 4 lines", "8 B buffer gather, 64 lines", "the same, sc0", "the same, nt",
             "the same, sc1", "the same, sc0 sc1", "the same, sc0 sc1 nt"]
    big = 1 << 18
    buf = hip.DeviceArray.from_host(1.0 + (np.arange(big) & 7).astype(np.float64))
    sink = hip.DeviceArray(1 << 17)
    out["load_cost_cycles_per_cu"] = {}
    for n, where in ((2048, "L1 (16 KiB)"), (big, "L2 (2 MiB)")):
        row = {}
        for mode, name in enumerate(names):
            row[name] = round(hip.gather_probe(mode, buf, n, 4096, sink), 1)
        out["load_cost_cycles_per_cu"][where] = row
        log("load cost, %s: %s" % (where, row))
    print(json.dumps({"probe": out}), flush=True)


def run_e2e(args, gen, hip, np):
    """End-to-end streaming of the held-out workload (master seed 55: rows =
    cols log-uniform in [1e4, 4e6], avg 2..128, gamma rows, skew up to 1000;
    the other descriptors as the batched workload): the first `count`
    matrices of that stream, in chunks that fit HBM, through the selector (with
    its fallback to exhaustive tuning) and one batched launch per chunk,
    against plain per-matrix CSR launches; every y checked against the device
    csr_naive. spmvdare/stream.py holds the pipeline. Prints one JSON line."""
    from spmvdare import stream, tuning
    model = os.path.join(ROOT, "spmv-dare_amd", "spmvdare", "models", "selector.json")
    sel = tuning.Selector.load(model)
    count = args.e2e_count
    params = tuning.sample_params(55, max(count * 2, count + 64), **tuning.STREAM_RANGES)
    draws = "first %d matrices" % count
    if args.e2e_draws:
        # a slice of the stream by draw index (every generatable matrix of
        # draws [a, b)), so that a long stream can be run in several parts
        a_, b_ = (int(v) for v in args.e2e_draws.split(":"))
        params = tuning.sample_params(55, b_, **tuning.STREAM_RANGES)[a_:b_]
        count = b_ - a_
        draws = "every generatable matrix of draws [%d, %d)" % (a_, b_)
    r = stream.run_stream(params, count, max(1, args.e2e_chunk), sel, tuned=args.e2e_tuned,
                          fallback_spread=args.fallback_spread, steps=args.e2e_steps, batch_bnnz=args.batch_bnnz,
                          queues=args.queues, tile=args.tile, log=log, refine=args.refine)
    out = {
        "metric": "end-to-end batched FP64 SpMV matrices/s (selector-chosen formats), streaming",
        "value": round(r["matrices_per_s"], 1), "unit": "matrices/s", "n_gpus": 1, "higher_is_better": True,
        "dtype": "f64", "data": "synthetic (reference generator restated bit-exactly)",
        "config": {"workload": "%s (%d matrices) of the master-seed-55 stream (rows 1e4..4e6, avg 2..128, gamma, "
                               "skew <= 1000), chunks of %d" % (draws, r["matrices"], r["chunk"]),
                   "nnz": r["nnz"], "alg_bytes": r["alg_bytes"]},
        "batched_ms": round(r["batched_ms"], 3), "batched_gbps": round(r["batched_gbps"], 1),
        "plain_per_matrix_csr": {"ms": round(r["plain_per_matrix_csr_ms"], 3),
                                 "matrices_per_s": round(r["plain_matrices_per_s"], 1)},
        "speedup_vs_plain_csr": round(r["speedup_vs_plain_csr"], 3),
        "max_normwise_err_vs_csr_naive": r["max_normwise_err_vs_csr_naive"],
        "selector": {"fallback_spread": r["fallback_spread"], "fallbacks_to_exhaustive": r["fallbacks"],
                     "refine_top_k": r["refine"],
                     "regret_vs_exhaustive": r["regret"]},
        "wall_s": round(r["wall_s"], 1), "chosen_variants": r["chosen_variants"],
    }
    log("e2e %d matrices: batched %.2f ms vs per-matrix CSR %.2f ms (x%.2f), max err %.2e, %d fallbacks" % (
        r["matrices"], r["batched_ms"], r["plain_per_matrix_csr_ms"], r["speedup_vs_plain_csr"],
        r["max_normwise_err_vs_csr_naive"], r["fallbacks"]))
    print(json.dumps(out), flush=True)


def batched_leg(args, gen, hip, np, count, steps, warmup):
    """Batched workload: `count` generated matrices of uneven sizes (rows =
    cols log-uniform in [1e3, 2e5], avg 4..64, ...; master seed 21), each with
    its own x, computed by one persistent work-queue launch replayed from a
    hipGraph, with every matrix's format chosen by the learned selector (a
    DIV-RF choice kept only when predicted `--batch-div-margin` faster than
    the best other variant: DIV members run in a second persistent kernel);
    compared on the same GPU with plain per-matrix binned-CSR launches (one
    launch per matrix, back to back) and with the same persistent launch over
    all-CSR members. Every member's batched y is checked against the device
    csr_naive of that matrix (a sequential left-to-right sum per row, the
    oracle's csr_naive bit for bit). Returns a dict."""
    from spmvdare import tuning
    model = os.path.join(ROOT, "spmv-dare_amd", "spmvdare", "models", "selector.json")
    sel = tuning.Selector.load(model) if os.path.exists(model) and not args.no_selector else None
    fixed = tuning.VARIANTS.index(("sell_c_s", {"sigma": 256, "lanes_per_row": 4, "unroll": 4}))
    if args.batch_fixed is not None:
        fixed, sel = args.batch_fixed, None
    csr = tuning.DEFAULT_CSR
    allowed = tuning.batch_allowed()
    params = tuning.sample_params(21, count * 2)
    t0 = time.perf_counter()
    members, total_nnz, total_bytes, chosen = [], 0, 0, []
    for p in params:
        if len(members) == count:
            break
        try:
            a = gen.generate(p, stats=False)
        except gen.GenError:
            continue
        dm = tuning.DeviceCSR(a, p.seed + 1)
        if sel is not None:
            v = tuning.batch_choice(sel.predict(tuning.features(p, dm.stats()))[0], allowed, args.batch_bnnz,
                                    args.batch_div_margin)
        else:
            v = fixed
        chosen.append(v)
        f_sel = dm.convert(v)
        f_csr = dm.convert(csr)
        dy_csr = hip.DeviceArray(max(a.m, 1))
        members.append((dm, f_sel, f_csr, dy_csr))
        total_nnz += a.nnz
        total_bytes += alg_bytes(a.m, a.n, a.nnz)
        dm.a = None  # host arrays no longer needed
        if len(members) % 512 == 0:
            log("batched: %d/%d matrices prepared, %.0f s" % (len(members), count, time.perf_counter() - t0))
    prep_s = time.perf_counter() - t0
    log("batched: %d matrices, %d nnz, prepared in %.0f s; selector=%s" % (
        len(members), total_nnz, prep_s, "mlp" if sel is not None else "fixed"))
    batch = hip.Batch(args.tile)
    batch.set_queues(args.queues)
    base = hip.Batch(args.tile)
    for dm, f_sel, f_csr, dy_csr in members:
        batch.add(f_sel, dm.dx, dm.dy)
        base.add(f_csr, dm.dx, dy_csr)
    batch.finalize(use_graph=not args.no_graph)
    base.finalize(use_graph=False)
    hip.synchronize()
    # The three legs are timed in interleaved rounds (each round: an untimed
    # iteration, then the timed ones, of every leg in turn; the plain
    # launches first, then the two persistent launches in alternating
    # order), so that they run at the same clock and cache state: timed one
    # after the other, the same launch ran about 2 % slower first than last.
    # The third leg is the same persistent launch with every member as plain
    # binned CSR: the batching gain without the selector.
    batch.time(warmup, 1)
    base.time(1, 1, per_matrix=True)
    rounds = 4
    ms_batch, ms_plain, ms_base = [], [], []
    for r in range(rounds):
        ms_plain += list(base.time(1, max(1, steps // (4 * rounds)), per_matrix=True))
        for leg in ((0, 1) if r % 2 == 0 else (1, 0)):
            if leg == 0:
                ms_batch += list(batch.time(1, max(1, steps // rounds)))
            else:
                ms_base += list(base.time(1, max(1, steps // rounds)))
    ms_batch, ms_plain, ms_base = np.array(ms_batch), np.array(ms_plain), np.array(ms_base)
    batch.run()
    batch.synchronize()
    # every member's batched y against its csr_naive y on the device
    t1 = time.perf_counter()
    worst, worst_csr = 0.0, 0.0
    for dm, f_sel, f_csr, dy_csr in members:
        yb, yc = dm.dy.download(), dy_csr.download()
        f_ref = hip.csr_to_format_device(dm.rp, dm.col, dm.val, dm.m, dm.n, dm.nnz, kind="csr_naive")
        f_ref.spmv_device(dm.dx, dy_csr)
        hip.synchronize()
        yr = dy_csr.download()
        f_ref.destroy()
        nrm = np.linalg.norm(yr)
        rel = (lambda y_: float(np.linalg.norm(y_ - yr) / nrm) if nrm > 0 else float(np.linalg.norm(y_)))
        worst, worst_csr = max(worst, rel(yb)), max(worst_csr, rel(yc))
    check_s = time.perf_counter() - t1
    t_b, t_p, t_a = float(np.median(ms_batch)), float(np.median(ms_plain)), float(np.median(ms_base))
    hist = np.bincount(np.array(chosen), minlength=len(tuning.VARIANTS))
    out = {
        "workload": "%d generated matrices, master seed 21 (rows 1e3..2e5, avg 4..64 nnz/row)" % len(members),
        "matrices": len(members), "nnz": total_nnz, "alg_bytes": total_bytes,
        "selector": "mlp" if sel is not None else tuning.variant_name(fixed),
        "steps": steps, "ms": round(t_b, 4), "min_ms": round(float(np.min(ms_batch)), 4),
        "matrices_per_s": round(len(members) / (t_b * 1e-3), 1),
        "gbps": round(total_bytes / (t_b * 1e-3) / 1e9, 1),
        "plain_per_matrix_csr": {"ms": round(t_p, 4), "matrices_per_s": round(len(members) / (t_p * 1e-3), 1),
                                 "gbps": round(total_bytes / (t_p * 1e-3) / 1e9, 1)},
        "speedup_vs_plain_csr": round(t_p / t_b, 3),
        "batched_all_csr": {"ms": round(t_a, 4), "speedup_vs_plain_csr": round(t_p / t_a, 3)},
        "max_normwise_err_vs_csr_naive": worst,
        "max_normwise_err_vs_csr_naive_plain_csr": worst_csr,
        "prep_s": round(prep_s, 1), "check_s": round(check_s, 1),
        "batch_min_block_nnz": args.batch_bnnz, "batch_div_margin": args.batch_div_margin,
        "hip_graph": not args.no_graph,
        "chosen_variants": {tuning.variant_name(i): int(c) for i, c in enumerate(hist) if c},
    }
    log("batched %.3f ms (%.0f matrices/s, %.0f GB/s) vs plain per-matrix CSR %.3f ms: x%.2f; "
        "batched all-CSR %.3f ms; max err vs csr_naive %.2e" % (
            t_b, out["matrices_per_s"], out["gbps"], t_p, t_p / t_b, t_a, worst))
    batch.destroy()
    base.destroy()
    return out


def run_member_times(args, gen, hip, np):
    """--member-times: every member of the batched workload timed alone as a
    plain binned-CSR launch (median of 10 after 3 warm-up launches, caches
    warm), summarised by member size and by neighbours per nonzero (how many
    of a row's nonzeros share cache lines of x). Prints one JSON line; it
    shows which members a faster gather would help."""
    from spmvdare import tuning
    recs = []
    for p in tuning.sample_params(21, args.count * 2):
        if len(recs) == args.count:
            break
        try:
            a = gen.generate(p, stats=False)
        except gen.GenError:
            continue
        dm = tuning.DeviceCSR(a, p.seed + 1)
        f = dm.convert(tuning.DEFAULT_CSR)
        ms = float(np.median(f.time(dm.dx, dm.dy, 3, 10)))
        f.destroy()
        recs.append((a.nnz, alg_bytes(a.m, a.n, a.nnz), p.avg_num_neighbours, p.placement, ms))
        del dm, a
        if len(recs) % 256 == 0:
            log("member times: %d/%d" % (len(recs), args.count))
    nnz = np.array([r[0] for r in recs], dtype=np.float64)
    byt = np.array([r[1] for r in recs], dtype=np.float64)
    nb = np.array([r[2] for r in recs])
    rnd = np.array([r[3] == "random" for r in recs])
    ms = np.array([r[4] for r in recs])

    def cls(mask):
        return {"members": int(mask.sum()), "ms": round(float(ms[mask].sum()), 3),
                "gbps": round(float(byt[mask].sum() / (ms[mask].sum() * 1e-3) / 1e9), 1) if mask.any() else None}

    out = {"workload": "%d generated matrices, master seed 21, each a plain binned-CSR launch" % len(recs),
           "sum_ms": round(float(ms.sum()), 3), "gbps": round(float(byt.sum() / (ms.sum() * 1e-3) / 1e9), 1),
           "by_nnz": {"%g..%g" % (lo, hi): cls((nnz >= lo) & (nnz < hi))
                      for lo, hi in ((0, 1e5), (1e5, 1e6), (1e6, 1e7), (1e7, 1e9))},
           "nnz_ge_1e6_by_neighbours": {"%g..%g" % (lo, hi): cls((nnz >= 1e6) & (nb >= lo) & (nb < hi))
                                        for lo, hi in ((0, 1), (1, 2), (2, 4), (4, 9))},
           "by_placement": {"random": cls(rnd), "diagonal": cls(~rnd)}}
    print(json.dumps(out), flush=True)


def run_batched(args, gen, hip, np):
    """--batched: the batched leg alone, as its own JSON line (matrices/s)."""
    b = batched_leg(args, gen, hip, np, args.count, args.steps, args.warmup)
    out = {"metric": "batched FP64 SpMV matrices/s (selector-chosen formats, one persistent hipGraph launch)",
           "value": b["matrices_per_s"], "unit": "matrices/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": b["ms"], "higher_is_better": True, "dtype": "f64",
           "data": "synthetic (reference generator restated bit-exactly)",
           "config": {"workload": b["workload"]}, "batched": b}
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--probe", action="store_true", help="measure stream read / copy bandwidth only")
    ap.add_argument("--batched", action="store_true", help="batched workload only (its own JSON line)")
    ap.add_argument("--member-times", action="store_true",
                    help="each member of the batched workload timed alone as plain CSR, by member class")
    ap.add_argument("--no-batched", action="store_true", help="leave the batched leg out of the default run")
    ap.add_argument("--count", type=int, default=2048, help="matrices of the batched workload")
    ap.add_argument("--tile", type=int, default=0,
                    help="batched launch: consecutive units per ticket (0 = the library default)")
    ap.add_argument("--no-selector", action="store_true")
    ap.add_argument("--no-graph", action="store_true",
                    help="batched launch: launch the persistent kernels directly instead of replaying a hipGraph")
    ap.add_argument("--batch-fixed", type=int, default=None,
                    help="batched launch: every member as this variant (index into the variant list), no selector")
    ap.add_argument("--batch-bnnz", type=int, default=256,
                    help="batched launch: CSR choices with smaller row blocks run with this block size (0 = as chosen)")
    ap.add_argument("--batch-div-margin", type=float, default=0.25,
                    help="batched launch: a DIV-RF choice is kept only when predicted this much faster than the "
                         "best other variant (DIV members run in a second persistent kernel)")
    ap.add_argument("--queues", type=int, default=256,
                    help="batched launch: ticket counters the units are dealt over (a power of two)")
    ap.add_argument("--e2e", action="store_true", help="end-to-end streaming sample of the held-out workload")
    ap.add_argument("--e2e-count", type=int, default=1000, help="matrices of the stream")
    ap.add_argument("--e2e-chunk", type=int, default=16, help="matrices per chunk of the stream")
    ap.add_argument("--e2e-tuned", type=int, default=32, help="matrices of the stream also tuned exhaustively")
    ap.add_argument("--fallback-spread", type=float, default=0.3,
                    help="selector falls back to exhaustive tuning when the timed ratios of its best predictions "
                         "differ from the predicted ratios by more than this")
    ap.add_argument("--e2e-steps", type=int, default=10, help="timed launches per chunk")
    ap.add_argument("--e2e-draws", default=None, help="A:B -- every generatable matrix of stream draws [A, B)")
    ap.add_argument("--refine", type=int, default=3, help="time the k best predicted variants, keep the fastest")
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=128)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--workload", default="spmv8m", choices=sorted(WORKLOADS))
    ap.add_argument("--formats", default="all",
                    help="comma list of kind[:sigma], or 'all'")
    ap.add_argument("--cpu-budget", type=float, default=10.0, help="seconds of the port's timing loop")
    ap.add_argument("--cpu-timeout", type=int, default=150, help="seconds per reference CPU format run")
    ap.add_argument("--no-cpu", action="store_true", help="no CPU baseline and no oracle check of y")
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU: every GPU box here has one MI355X and this "
              "path does not shard across GPUs (--gpus %d requested)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import numpy as np
    from spmvdare import gen, hip

    hip.init(0)
    log("device:", hip.device_name())
    if args.probe:
        probe(hip, np)
        return
    if args.batched:
        run_batched(args, gen, hip, np)
        return
    if args.e2e:
        run_e2e(args, gen, hip, np)
        return
    if args.member_times:
        run_member_times(args, gen, hip, np)
        return
    line, desc = WORKLOADS[args.workload]
    p = gen.GenParams.parse(line)
    t0 = time.perf_counter()
    a = gen.generate(p, stats=False)
    log("generated %s: m=%d nnz=%d in %.1f s" % (args.workload, a.m, a.nnz, time.perf_counter() - t0))
    x = gen.make_x(a.n, p.seed + 1)
    B = alg_bytes(a.m, a.n, a.nnz)
    m, n, nnz = a.m, a.n, a.nnz

    cands = CANDIDATES
    if args.formats == "sweep":
        from spmvdare import tuning
        cands = list(tuning.VARIANTS) + [("csr_naive", {})]
    elif args.formats != "all":
        cands = [parse_format(tok) for tok in args.formats.split(",")]

    dx = hip.DeviceArray.from_host(x)
    dy = hip.DeviceArray(a.m)
    results, ys = [], {}
    best = None
    for kind, kw in cands:
        t0 = time.perf_counter()
        f = hip.csr_to_format(a.row_ptr, a.col_ind, a.values, a.m, a.n, a.nnz, kind=kind, **kw)
        info = f.info()
        hip.synchronize()
        t_conv = time.perf_counter() - t0
        steps = args.steps if kind != "csr_naive" else min(args.steps, 16)
        dy.fill_bytes(0xff)  # NaN: a row the kernel does not write shows up in the check
        ms = f.time(dx, dy, args.warmup, steps)
        hip.synchronize()
        med = float(np.median(ms))
        r = {"format": kind, "params": kw, "median_ms": med, "min_ms": float(ms.min()),
             "mean_ms": float(ms.mean()), "gbps": B / (med * 1e-3) / 1e9,
             "gflops": 2 * a.nnz / (med * 1e-3) / 1e9, "conv_ms_device": info["conv_ms"],
             "csr_to_format_s": t_conv, "padded_nnz": info["padded_nnz"],
             "mem_footprint": info["mem_footprint"], "steps": steps}
        key = "%s %s" % (kind, json.dumps(kw, sort_keys=True))
        r["key"] = key
        ys[key] = dy.download()
        results.append(r)
        log("%-10s %-16s median %8.4f ms (min %.4f, mean %.4f)  %7.1f GB/s  %5.1f%% peak  conv %.1f ms" % (
            kind, kw, med, r["min_ms"], r["mean_ms"], r["gbps"], 100 * r["gbps"] / HBM_PEAK_GBS, info["conv_ms"]))
        if best is None or r["gbps"] > best["gbps"]:
            best = r
        f.destroy()
    r = best
    traffic = profiled_traffic(args.workload, r["format"], r["params"])

    # cpu leg: the oracle as the checker of every y, then the CPU baseline
    cpu, errs, dd, dd_naive = None, None, None, None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        errs, dd, dd_naive = check_leg(a, x, ys, r["key"])
        for res in results:
            res["normwise_err_vs_csr_naive"] = errs[res["key"]]
        ys.clear()
        cpus, _, _ = host_cpus()
        log("cpu baseline on %d physical cores ..." % len(cpus))
        cpu = cpu_baseline_reference(line, a.m, a.n, a.nnz, args.cpu_timeout)
        if cpu is None:
            cpu = cpu_baseline_port(a, x, args.cpu_budget)
            log("cpu baseline (port) %.1f GB/s" % cpu["value"])
    ys.clear()

    batched = None
    if not args.no_batched:
        del a
        batched = batched_leg(args, gen, hip, np, args.count, args.steps, args.warmup)

    out = {
        "metric": "FP64 SpMV GB/s (% HBM peak); learned-selected batched vs plain CSR",
        "value": round(r["gbps"], 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": r["steps"],
        "warmup": args.warmup,
        "ms_per_step": round(r["median_ms"], 6),
        "higher_is_better": True,
        "scaling": "strong",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic: reference generator restated bit-exactly; values 2v-1, x from seed+1",
        "config": {"workload": desc, "m": m, "n": n, "nnz": nnz,
                   "format": r["format"], "params": r["params"],
                   "pct_hbm_peak": round(100 * r["gbps"] / HBM_PEAK_GBS, 2),
                   "gflops": round(r["gflops"], 2), "alg_bytes_per_step": B, "timing": "median of the timed steps",
                   "min_ms": round(r["min_ms"], 6), "mean_ms": round(r["mean_ms"], 6),
                   "formats": [{k: v for k, v in x_.items()
                                if k in ("format", "params", "median_ms", "min_ms", "gbps", "normwise_err_vs_csr_naive")}
                               for x_ in results]},
        "normwise_err_vs_csr_naive": errs[r["key"]] if errs else None,
        "normwise_err_vs_double_double": dd,
        "roofline": {"bound": "hbm", "achieved": round(B / (r["median_ms"] * 1e-3) / 1e9, 2),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(B / (r["median_ms"] * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                     "traffic": traffic},
        "cpu_baseline": cpu,
        "batched": batched,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
