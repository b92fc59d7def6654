"""Autotuner and learned format selector.

Variant space: 40 format x tuning choices over the batchable drop-in formats
(row-length-binned CSR with rows-per-wave, lanes per row, load width and XCD
chunks; SELL with lanes per row and its inner unroll; SELL-C-sigma with
sigma, lanes per row, unroll and x windows; SELL_SORTED; DIV-RF with packet
size, values per decode step, device layout and packet reordering). The exhaustive
autotuner converts a matrix to every variant on the device and times it; the
selector is an FP32 MLP 14 -> 64 -> 64 -> 40 (hand-written HIP kernels,
spmvhip_mlp_*) that predicts every variant's slowdown log(t / t_best) from
the 10 generator descriptors plus 4 device-computed structure statistics,
trained on the autotuner's own GPU timings. It only changes which kernel
runs; its regret is judged against exhaustive tuning, and with `refine` the
top-k predicted variants are timed and the fastest kept (the fallback).
Nothing here has a counterpart in the reference.
"""
import hashlib
import json
import math
import os
import random

import numpy as np

from . import gen, hip


def _variants():
    v = []
    # binned CSR: rows per wave through the row-block size, lanes per row in the reduction
    for bnnz in (64, 128, 256, 512):
        for lanes in (0, 1):
            v.append(("csr_vec", {"block_nnz": bnnz, "lanes_per_row": lanes}))
    for bnnz in (128, 256):
        v.append(("csr_vec", {"block_nnz": bnnz, "lanes_per_row": 4}))
    # one stored entry per lane and load; row blocks dealt to the XCDs in chunks
    v.append(("csr_vec", {"block_nnz": 256, "lanes_per_row": 0, "narrow_loads": 1}))
    v.append(("csr_vec", {"block_nnz": 256, "lanes_per_row": 0, "xcd_remap": 1}))
    v.append(("csr_vec", {"block_nnz": 512, "lanes_per_row": 0, "xcd_remap": 1}))
    v.append(("csr_vec", {"block_nnz": 256, "lanes_per_row": 4, "xcd_remap": 1}))
    # SELL-C-sigma, C = 64: sort window sigma x lanes per row T
    for sigma in (64, 256, 1024, 4096, 16384):
        v.append(("sell_c_s", {"sigma": sigma, "lanes_per_row": 1, "unroll": 4}))
    for sigma in (64, 256, 4096):
        v.append(("sell_c_s", {"sigma": sigma, "lanes_per_row": 2, "unroll": 4}))
    v.append(("sell_c_s", {"sigma": 256, "lanes_per_row": 4, "unroll": 4}))
    for unroll in (2, 8):
        v.append(("sell_c_s", {"sigma": 256, "lanes_per_row": 1, "unroll": unroll}))
    v.append(("sell_c_s", {"sigma": 4096, "lanes_per_row": 4, "unroll": 4, "y_mode": 3}))
    # x windows in LDS (banded placements)
    v.append(("sell_c_s", {"sigma": 256, "lanes_per_row": 1, "unroll": 4, "x_window": 1}))
    v.append(("sell_c_s", {"sigma": 64, "lanes_per_row": 4, "unroll": 4, "x_window": 4096}))
    # unsorted SELL
    for T in (1, 2):
        v.append(("sell", {"lanes_per_row": T, "unroll": 4}))
    v.append(("sell", {"lanes_per_row": 1, "unroll": 8}))
    # DIV-RF: packet size P x values per lane and decode step, automatic device layout
    # (interleaved lane streams), one encoder partition, groups dealt to the XCDs in chunks
    for P, unroll in ((1024, 8), (1024, 4), (512, 4), (512, 8), (2048, 8)):
        v.append(("div_rf", {"div_packet_vals": P, "div_threads": 1, "unroll": unroll, "xcd_remap": 1}))
    v.append(("div_rf", {"div_packet_vals": 1024, "div_threads": 1, "unroll": 8, "xcd_remap": 1, "div_layout": 1}))
    # SELL_SORTED (ascending sort in row windows) and DIV-RF with the packet groups in median-column order
    v.append(("sell_sorted", {"sigma": 4096, "lanes_per_row": 1, "unroll": 4}))
    v.append(("sell_sorted", {"sigma": 256, "lanes_per_row": 2, "unroll": 4}))
    v.append(("div_rf", {"div_packet_vals": 1024, "div_threads": 1, "unroll": 8, "xcd_remap": 1, "div_reorder": 1}))
    return v


VARIANTS = _variants()
assert len(VARIANTS) == 40
# the plain per-matrix CSR launch the batched path is compared with
DEFAULT_CSR = VARIANTS.index(("csr_vec", {"block_nnz": 256, "lanes_per_row": 0}))


def batch_allowed():
    """Variants a member of the persistent batched launch may use: the launch
    stages CSR row blocks of at most 256 nonzeros at full occupancy (a
    512-nonzero member would halve it for every member)."""
    return np.array([not (k == "csr_vec" and kw.get("block_nnz", 256) > 256) for k, kw in VARIANTS])


def variant_name(i):
    k, p = VARIANTS[i]
    return k + "(" + ",".join("%s=%s" % kv for kv in sorted(p.items())) + ")"


# ------------------------------------------------------------------ workloads

def sample_params(master_seed, count, rows=(1e3, 2e5), avg=(4, 64), std_ratio=(0.25, 2.0),
                  bw=(0.01, 1.0), skew=(1, 100), neigh=(0, 8), crs=(0, 0.9),
                  distributions=("normal", "gamma"), placements=("random", "diagonal")):
    """Generator parameter sets drawn from ranges (rows = cols log-uniform);
    the batched workloads' sampling rule."""
    rnd = random.Random(master_seed)
    out = []
    for _ in range(count):
        r = int(round(10 ** rnd.uniform(math.log10(rows[0]), math.log10(rows[1]))))
        a = rnd.uniform(*avg)
        s = a * rnd.uniform(*std_ratio)
        d = rnd.choice(distributions)
        pl = rnd.choice(placements)
        b = rnd.uniform(*bw)
        sk = rnd.uniform(*skew)
        ng = rnd.uniform(*neigh)
        cr = rnd.uniform(*crs)
        seed = rnd.randrange(1, 2 ** 31)
        out.append(gen.GenParams.parse("%d %d %.6g %.6g %s %s %.6g %.6g %.6g %.6g %d" % (
            r, r, a, s, d, pl, b, sk, ng, cr, seed)))
    return out


# The streaming workload's ranges (rows = cols log-uniform in [1e4, 4e6], avg
# 2..128, gamma rows, skew up to 1000; the other descriptors as above).
STREAM_RANGES = dict(rows=(1e4, 4e6), avg=(2, 128), skew=(1, 1000), distributions=("gamma",))
# sample_params keyword ranges by workload name
WORKLOAD_RANGES = {"batched": {}, "stream": STREAM_RANGES}


# ------------------------------------------------------------------ features

FEATURE_NAMES = ["log_rows", "log_cols", "log_avg", "std_over_avg", "gamma", "diagonal", "bw", "log1p_skew",
                 "neigh", "crs", "cv_rows", "max_over_avg", "bw_measured", "neigh_per_nnz"]


def features(p, st):
    """14 features: the 10 generator descriptors (log-scaled where they span
    decades) and 4 device-computed structure statistics."""
    avg = max(st["avg_nnz_per_row"], 1e-9)
    return np.array([
        math.log(p.nr_rows), math.log(p.nr_cols), math.log(max(p.avg_nnz_per_row, 1e-3)),
        p.std_nnz_per_row / max(p.avg_nnz_per_row, 1e-3),
        1.0 if p.distribution.startswith("g") else 0.0, 1.0 if p.placement.startswith("d") else 0.0,
        min(p.bw, 1.0), math.log1p(max(p.skew, 0.0)), p.avg_num_neighbours, p.cross_row_similarity,
        st["std_nnz_per_row"] / avg, st["max_nnz_per_row"] / avg, st["avg_bw_scaled"],
        st["avg_num_neighbours"]], dtype=np.float64)


class DeviceCSR:
    """A generated matrix resident on the device, with its x and y."""

    def __init__(self, a, x_seed):
        self.a = a
        self.m, self.n, self.nnz = a.m, a.n, a.nnz
        self.rp = hip.DeviceArray.from_host(a.row_ptr)
        self.col = hip.DeviceArray.from_host(a.col_ind if a.nnz else np.zeros(1, np.int32))
        self.val = hip.DeviceArray.from_host(a.values if a.nnz else np.zeros(1))
        self.x = gen.make_x(a.n, x_seed)
        self.dx = hip.DeviceArray.from_host(self.x)
        self.dy = hip.DeviceArray(max(a.m, 1))

    def stats(self):
        return hip.csr_stats_device(self.rp, self.col, self.m, self.n)

    def convert(self, i):
        kind, kw = VARIANTS[i]
        return hip.csr_to_format_device(self.rp, self.col, self.val, self.m, self.n, self.nnz, kind=kind, **kw)


def batch_variant(i, min_block_nnz=256):
    """The variant a matrix runs as inside a persistent batched launch.

    The selector and the autotuner rank variants by standalone launches,
    where a small matrix gains from small CSR row blocks, one entry per lane
    and load, or a fixed reduction width (more waves in flight, fewer
    idle lanes). In a batch the other members supply that parallelism, and
    each of these only adds per-unit work to a wave's sequential stream:
    every member as one variant, 2,048 matrices, on one stream, measured the
    default (256-nonzero blocks, paired loads, automatic reduction width)
    best, narrow loads 2 % and one lane per row 1 % slower. So a CSR choice
    runs as that default, with row blocks of max(256, min_block_nnz) if its
    own are smaller (min_block_nnz 0 keeps a choice's block size). Other
    formats are kept.
    """
    kind, kw = VARIANTS[i]
    if kind != "csr_vec":
        return i
    bnnz = kw.get("block_nnz", 256)
    want = {"block_nnz": max(bnnz, min_block_nnz) if min_block_nnz else bnnz, "lanes_per_row": 0}
    for j, (k2, kw2) in enumerate(VARIANTS):
        if k2 == kind and kw2 == want:
            return j
    return i


def batch_choice(pred, allowed, min_block_nnz=256, div_margin=0.25):
    """The variant a matrix runs as inside a persistent batched launch, from
    the selector's predicted slowdowns `pred` (log t_v / t_best, one row).

    DIV-RF members run in a second persistent kernel after the one that
    takes every other member, so DIV members add that kernel's launch, ramp
    and tail (on 2,048 small matrices, two DIV members cost the launch about
    1 %); a DIV choice is kept only when it is predicted at least
    `div_margin` faster than the best other allowed variant (else that
    variant runs). The choice then goes through batch_variant."""
    pred = np.where(np.asarray(allowed, dtype=bool), np.asarray(pred, dtype=np.float64), np.inf)
    best = int(np.argmin(pred))
    if VARIANTS[best][0] == "div_rf":
        other = np.where(np.array([k != "div_rf" for k, _ in VARIANTS]), pred, np.inf)
        alt = int(np.argmin(other))
        if np.isfinite(other[alt]) and other[alt] - pred[best] < math.log1p(div_margin):
            best = alt
    return batch_variant(best, min_block_nnz)


def time_variant(dm, i, warmup=2, iters=8):
    f = dm.convert(i)
    try:
        ms = f.time(dm.dx, dm.dy, warmup, iters)
    finally:
        f.destroy()
    return float(np.median(ms))


def autotune(dm, variants=None, warmup=2, iters=8, cache=None, key=None):
    """Exhaustive: median time (ms) of every variant. With an AutotuneCache
    and the matrix's key, variants timed before on this device are read from
    the cache and the new timings are added to it."""
    idx = list(range(len(VARIANTS)) if variants is None else variants)
    known = cache.get(key) if cache is not None and key is not None else {}
    out = []
    for i in idx:
        name = variant_name(i)
        if name not in known:
            known[name] = time_variant(dm, i, warmup, iters)
        out.append(known[name])
    if cache is not None and key is not None:
        cache.put(key, known)
    return np.array(out)


class AutotuneCache:
    """Autotuner timings kept on disk (JSON): per device name and matrix key
    (matrix_key), the median ms of every variant timed so far. Written back
    on every put, through a temporary file, so an interrupted tuning run
    resumes where it stopped."""

    def __init__(self, path, device=None):
        self.path = path
        self.device = device or hip.device_name()
        try:
            with open(path) as f:
                self.data = json.load(f)
        except (OSError, ValueError):
            self.data = {}

    def get(self, key):
        return dict(self.data.get(self.device, {}).get(key, {}))

    def put(self, key, timings):
        self.data.setdefault(self.device, {})[key] = dict(timings)
        tmp = self.path + ".tmp"
        with open(tmp, "w") as f:
            json.dump(self.data, f)
        os.replace(tmp, self.path)


def matrix_key(a, params=None):
    """Cache key of a matrix: its generator parameter line when it was
    generated (the 11 parameters fix it), else a digest of its arrays."""
    if params is not None:
        return "gen:" + params.line()
    h = hashlib.sha256()
    h.update(np.array([a.m, a.n, a.nnz], dtype=np.int64).tobytes())
    for arr in (a.row_ptr, a.col_ind, a.values):
        h.update(np.ascontiguousarray(arr).tobytes())
    return "sha256:" + h.hexdigest()


# ------------------------------------------------------------------ selector

class Selector:
    """MLP over standardised features predicting log(t_variant / t_best)."""

    def __init__(self, hidden=(64, 64), seed=0):
        self.hidden = hidden
        self.seed = seed
        self.mu = None
        self.sd = None
        self.net = None

    def _norm(self, X):
        return (np.asarray(X, dtype=np.float64) - self.mu) / self.sd

    @staticmethod
    def targets(times):
        t = np.asarray(times, dtype=np.float64)
        return np.log(t / t.min(axis=1, keepdims=True))

    def fit(self, X, times, epochs=300, batch=128, lr=2e-3, shuffle_seed=0, resume=False):
        """Train on (features, timings). resume: continue from the current
        weights, feature scaling and Adam state (a loaded checkpoint) instead
        of a fresh network."""
        X = np.asarray(X, dtype=np.float64)
        T = self.targets(times)
        if not (resume and self.net is not None):
            self.mu = X.mean(axis=0)
            self.sd = X.std(axis=0) + 1e-9
            self.net = hip.MLP(X.shape[1], self.hidden[0], self.hidden[1], T.shape[1], seed=self.seed)
        rng = np.random.default_rng(shuffle_seed)
        Xn = self._norm(X).astype(np.float32)
        losses = []
        for _ in range(epochs):
            order = rng.permutation(len(X))
            losses.append(float(self.net.train(Xn[order], T[order].astype(np.float32), batch=batch,
                                               epochs=1, lr=lr)[0]))
        return np.array(losses)

    def predict(self, X):
        return self.net.predict(self._norm(np.atleast_2d(X)).astype(np.float32))

    def select(self, X, k=1, allowed=None):
        """Indices of the k best predicted variants per row (only among the
        variants where the boolean mask `allowed` is true, if given)."""
        pred = self.predict(X)
        if allowed is not None:
            pred = np.where(np.asarray(allowed, dtype=bool)[None, :], pred, np.inf)
        return np.argsort(pred, axis=1)[:, :k]

    def save(self, path, adam=False):
        """Weights, feature scaling and variant list as JSON; with adam, also
        Adam's moments and step (a checkpoint that fit(resume=True) continues)."""
        d = {"hidden": list(self.hidden), "mu": self.mu.tolist(), "sd": self.sd.tolist(),
             "dims": list(self.net.dims), "params": self.net.get_params().tolist(),
             "variants": [variant_name(i) for i in range(len(VARIANTS))]}
        if adam:
            m1, m2, step = self.net.get_adam_state()
            d["adam"] = {"m1": m1.tolist(), "m2": m2.tolist(), "step": int(step)}
        with open(path, "w") as f:
            json.dump(d, f)

    @classmethod
    def load(cls, path):
        """A trained selector. Its output slots are variant indices, so the
        variant list it was trained on must be the current VARIANTS."""
        d = json.load(open(path))
        have = [variant_name(i) for i in range(len(VARIANTS))]
        if d.get("variants") != have:
            raise ValueError("%s was trained on another variant list; retrain it "
                             "(python -m spmvdare.train_selector)" % path)
        s = cls(tuple(d["hidden"]))
        s.mu = np.array(d["mu"])
        s.sd = np.array(d["sd"])
        s.net = hip.MLP(*d["dims"])
        s.net.set_params(np.array(d["params"], dtype=np.float32))
        if "adam" in d:
            s.net.set_adam_state(np.array(d["adam"]["m1"], dtype=np.float32),
                                 np.array(d["adam"]["m2"], dtype=np.float32), d["adam"]["step"])
        return s


def regret(times_row, choice):
    return float(times_row[choice] / np.min(times_row) - 1.0)
