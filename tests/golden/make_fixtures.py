"""Write the golden fixtures from the reference's own code.

Runs in the development container only (it needs the reference sources,
compiled into oracle/_ref/ by `make -C oracle ref`). The outputs are data:
hashes, descriptor statistics and a few small arrays produced by the
reference's generator (artificial-matrix-generator/artificial_matrix_generation.c)
and SELL-C-sigma conversion (src/spmv_kernels/sell-C-s/radix_sort.c,
sellcs_format.c).

    python tests/golden/make_fixtures.py [sell|sell_sorted|div_reorder|div_lossy]   (one set only)

Outputs (tests/golden/):
    generator.json  per parameter line: m, n, nnz, FNV-1a 64 of row_ptr /
                    col_ind / raw values, and the reference's descriptor stats
    sellcs.json     per (matrix, C, sigma): FNV-1a 64 of each converted array
    sell.json       per (matrix, threads): the reference's SELL
                    (src/spmv_kernels/sell.cpp, 4-row clusters): digests of its
                    cluster sizes, ja, a and y (x as for sell_sorted.json)
    sell_sorted.json  per (matrix, threads): the reference's SELL_SORTED
                    (src/spmv_kernels/sell_sorted.cpp, AVX2 build: 4-row
                    clusters): thread windows and digests of its permutation,
                    cluster sizes, ja, a and of y for x_j = ((j % 13) + 1)/8 - 1
    div_lossy.json  per (matrix, T, P, tolerance): the reference's DIV-RF
                    stream with DIV_VC_TOLERANCE (digest, header padding
                    masked); tiny_div_lossy.npz: one such stream in full
    div_reorder.json  per (matrix, T, P): the reference's DIV-RF stream with
                    DIV_PACKET_REORDERING=1, as the sequence of its packet
                    groups (row_min of each group's first packet) per partition
    tiny_csr.npz, tiny_sellcs.npz   full arrays of the 1000 x 1000 test matrix
"""
import json
import os
import random
import struct
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF = os.path.join(ROOT, "oracle", "_ref")

FIELDS = ["m", "n", "nnz", "h_row_ptr", "h_col_ind", "h_values", "avg_nnz_per_row",
          "std_nnz_per_row", "min_nnz_per_row", "max_nnz_per_row", "skew", "avg_bw",
          "avg_num_neighbours", "cross_row_similarity"]


def fixed_lines():
    return [
        # matrix_generation_parameters/test.txt:1
        ("test_txt", "1000 1000 5 1.6667 normal random 0.05 0 0.05 0.05 14"),
        # artificial-matrix-generator/run.sh:12-16
        ("run_sh_1", "65535 65535 5 1.6667 normal random 0.05 0 0.05 0.05 14"),
        ("run_sh_2", "1082401 1082401 10 3.3333 normal random 0.3 10000 0.5 0.5 14"),
        ("run_sh_3", "3020732 3020732 50 16.6667 normal random 0.3 100 1.9 0.05 14"),
        # the benchmark's CPU-runnable and large single-matrix workloads
        ("cfg1", "100000 100000 16 4 normal random 0.3 1.5 1 0.2 14"),
        ("cfg2", "8388608 8388608 32 48 gamma diagonal 0.05 20 4 0.5 7"),
        # edge cases of the generator's branches
        ("simple_placement", "5000 3000 12 6 normal simple 0.2 2 1 0.3 5"),
        ("bw_above_one", "4000 4000 20 5 gamma random 1.7 3 0.5 0.1 6"),
        ("std_zero", "3000 3000 9 0 normal diagonal 0.1 0.5 0 0 7"),
        ("skew_below_one", "3000 5000 9 3 gamma diagonal 0.02 0.5 1.5 0.9 8"),
        ("degree_clamp", "2000 40 30 25 gamma random 1 5 2 0.5 9"),
        ("min_rows", "100 100 4 2 normal random 0.5 1 1 0.5 10"),
        ("dense_rows", "300 300 250 40 normal random 1 0.2 1 0.5 11"),
        ("wide", "1000 200000 40 30 gamma diagonal 0.001 30 3 0.7 12"),
        ("tall", "200000 1000 3 2 gamma random 0.5 10 0 0 13"),
        ("neigh_two", "20000 20000 24 30 gamma diagonal 0.05 20 4 0.5 7"),
    ]


def sampled_lines(master_seed, count, rows_lo, rows_hi, avg_lo, avg_hi, skew_hi, tag):
    """Parameter sets drawn from the batched workloads' ranges."""
    rnd = random.Random(master_seed)
    out = []
    for k in range(count):
        rows = int(round(10 ** rnd.uniform(np.log10(rows_lo), np.log10(rows_hi))))
        avg = rnd.uniform(avg_lo, avg_hi)
        std = avg * rnd.uniform(0.25, 2.0)
        distr = rnd.choice(["normal", "gamma"])
        place = rnd.choice(["random", "diagonal"])
        bw = rnd.uniform(0.01, 1.0)
        skew = rnd.uniform(1, skew_hi)
        neigh = rnd.uniform(0, 8)
        crs = rnd.uniform(0, 0.9)
        seed = rnd.randrange(1, 2 ** 31)
        line = "%d %d %.6g %.6g %s %s %.6g %.6g %.6g %.6g %d" % (
            rows, rows, avg, std, distr, place, bw, skew, neigh, crs, seed)
        out.append(("%s_%02d" % (tag, k), line))
    return out


def run_gen(line, dump=None):
    args = [os.path.join(REF, "gen_dump")] + line.split() + ([dump] if dump else [])
    out = subprocess.run(args, check=True, capture_output=True, text=True).stdout.split()
    rec = {}
    for k, v in zip(FIELDS, out):
        rec[k] = int(v) if k in ("m", "n", "nnz") else (v if k.startswith("h_") else float(v))
    return rec


def read_sellcs(path, m):
    with open(path, "rb") as f:
        ns, total = struct.unpack("<qq", f.read(16))
        order = np.frombuffer(f.read(8 * m), dtype=np.int64)
        widths = np.frombuffer(f.read(8 * ns), dtype=np.int64)
        sptr = np.frombuffer(f.read(8 * (ns + 1)), dtype=np.int64)
        cols = np.frombuffer(f.read(8 * total), dtype=np.int64)
        vals = np.frombuffer(f.read(8 * total), dtype=np.float64)
    return order, widths, sptr, cols, vals


def read_csr(path):
    with open(path, "rb") as f:
        m, n, nnz = struct.unpack("<qqq", f.read(24))
        rp = np.frombuffer(f.read(4 * (m + 1)), dtype=np.int32)
        ci = np.frombuffer(f.read(4 * nnz), dtype=np.int32)
        va = np.frombuffer(f.read(8 * nnz), dtype=np.float64)
    return rp, ci, va


def digest(a):
    """Short content digest used by the sellcs fixtures (sha256 of bytes)."""
    import hashlib
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()[:32]


def main():
    if not os.path.exists(os.path.join(REF, "gen_dump")):
        sys.exit("oracle/_ref/gen_dump missing: run `make -C oracle ref` first")
    lines = fixed_lines()
    lines += sampled_lines(21, 24, 1e3, 2e5, 4, 64, 100, "batch")
    lines += sampled_lines(55, 8, 1e4, 4e5, 2, 128, 1000, "stream")
    gen = {}
    for name, line in lines:
        rec = run_gen(line)
        rec["line"] = line
        gen[name] = rec
        print(name, rec["nnz"], flush=True)
    with open(os.path.join(HERE, "generator.json"), "w") as f:
        json.dump(gen, f, indent=1, sort_keys=True)

    sell = {}
    with tempfile.TemporaryDirectory() as td:
        # sigma = 1 (< C) is left out: the reference writes past its arrays there
        # (sellcs_format.c:152 takes the first row of an unsorted slice as its
        # width) and glibc aborts the run.
        cases = [("test_txt", 64, 64), ("test_txt", 64, 256),
                 ("test_txt", 64, 4096), ("test_txt", 256, 16384),
                 ("cfg1", 64, 256), ("cfg1", 64, 4096), ("cfg1", 256, 16384),
                 ("neigh_two", 64, 256), ("neigh_two", 64, 4096),
                 ("batch_03", 64, 256), ("batch_07", 64, 4096), ("stream_01", 64, 4096),
                 ("tall", 64, 128), ("dense_rows", 64, 64)]
        dumped = {}
        for name, C, sigma in cases:
            line = gen[name]["line"]
            if name not in dumped:
                p = os.path.join(td, name + ".bin")
                run_gen(line, p)
                dumped[name] = p
            out = os.path.join(td, "s.bin")
            subprocess.run([os.path.join(REF, "sellcs_dump"), dumped[name], str(C), str(sigma), out],
                           check=True, capture_output=True)
            m = gen[name]["m"]
            order, widths, sptr, cols, vals = read_sellcs(out, m)
            key = "%s/C%d/s%d" % (name, C, sigma)
            sell[key] = {"line": line, "C": C, "sigma": sigma, "total": int(sptr[-1]),
                         "order": digest(order), "widths": digest(widths), "slice_ptr": digest(sptr),
                         "cols": digest(cols), "vals": digest(vals)}
            print(key, int(sptr[-1]), flush=True)
            if name == "test_txt" and C == 64 and sigma == 256:
                np.savez_compressed(os.path.join(HERE, "tiny_sellcs.npz"), order=order, widths=widths,
                                    slice_ptr=sptr, cols=cols, vals=vals)
        rp, ci, va = read_csr(dumped["test_txt"])
        np.savez_compressed(os.path.join(HERE, "tiny_csr.npz"), row_ptr=rp, col_ind=ci, values=va)
    with open(os.path.join(HERE, "sellcs.json"), "w") as f:
        json.dump(sell, f, indent=1, sort_keys=True)
    div_fixtures(gen)
    div_reorder_fixtures(gen)
    div_lossy_fixtures(gen)
    sell_fixtures(gen)
    sell_sorted_fixtures(gen)


def read_div(path):
    """Per partition: (stream bytes, packet group byte positions)."""
    with open(path, "rb") as f:
        T = struct.unpack("<q", f.read(8))[0]
        parts = []
        for _ in range(T):
            nb = struct.unpack("<q", f.read(8))[0]
            b = np.frombuffer(f.read(nb), dtype=np.uint8)
            ng = struct.unpack("<q", f.read(8))[0]
            pos = np.frombuffer(f.read(8 * (ng + 1)), dtype=np.int64)
            parts.append((b, pos))
    return parts


def mask_padding(b, L=4):
    """Zero the two padding bytes of every packet header (offsets 22-23),
    which the reference leaves uninitialised."""
    b = np.array(b, dtype=np.uint8, copy=True)
    at = 0
    while at < b.size:
        h = b[at:at + 28 + 4 * L].view(np.uint32)
        nv, nu, nrfs = int(h[0]), int(h[1]), int(h[6])
        cb = (int(b[at + 20]) + int(b[at + 21]) + 7) >> 3
        size = 28 + 4 * L + 8 * nrfs + nv * cb + nu + int(h[7:7 + L].astype(np.int64).sum())
        b[at + 22:at + 24] = 0
        at += size
    return b


def div_fixtures(gen):
    """DIV-RF streams of the reference (AVX2 build: 4 value lanes) at several
    partition counts T (OMP threads) and packet sizes P."""
    if not os.path.exists(os.path.join(REF, "div_dump")):
        sys.exit("oracle/_ref/div_dump missing: run `make -C oracle ref` first")
    cases = [("test_txt", 1, 16384), ("test_txt", 4, 512), ("test_txt", 3, 97),
             ("cfg1", 1, 16384), ("cfg1", 8, 4096), ("neigh_two", 2, 16384),
             ("batch_03", 5, 1000), ("dense_rows", 1, 2048), ("tall", 4, 16384),
             ("simple_placement", 2, 300), ("degree_clamp", 3, 16384)]
    out = {}
    with tempfile.TemporaryDirectory() as td:
        for name, T, P in cases:
            line = gen[name]["line"]
            csr = os.path.join(td, "c.bin")
            run_gen(line, csr)
            dump = os.path.join(td, "d.bin")
            env = dict(os.environ, OMP_NUM_THREADS=str(T))
            subprocess.run([os.path.join(REF, "div_dump"), csr, str(P), dump], check=True,
                           capture_output=True, env=env)
            parts = read_div(dump)
            stream = np.concatenate([mask_padding(b) for b, _ in parts]) if parts else np.zeros(0, np.uint8)
            groups = [int(p.size - 1) for _, p in parts]
            key = "%s/T%d/P%d" % (name, T, P)
            out[key] = {"line": line, "T": T, "L": 4, "P": P, "bytes": int(stream.size),
                        "stream": digest(stream), "groups_per_partition": groups,
                        "group_pos": digest(np.concatenate([p for _, p in parts]))}
            print(key, stream.size, flush=True)
            if key == "test_txt/T4/P512":
                np.savez_compressed(os.path.join(HERE, "tiny_div.npz"), stream=stream,
                                    **{"group_pos_%d" % t: p for t, (_, p) in enumerate(parts)})
    with open(os.path.join(HERE, "div.json"), "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)


def div_lossy_fixtures(gen):
    if not os.path.exists(os.path.join(REF, "div_dump")):
        sys.exit("oracle/_ref/div_dump missing: run `make -C oracle ref` first")
    cases = [("test_txt", 4, 512, "1e-6"), ("cfg1", 8, 4096, "1e-3"), ("batch_03", 5, 1000, "1e-9"),
             ("neigh_two", 2, 16384, "0.01"), ("dense_rows", 1, 2048, "1e-4"), ("tall", 3, 16384, "0.3")]
    out = {}
    with tempfile.TemporaryDirectory() as td:
        for name, T, P, tol in cases:
            line = gen[name]["line"]
            csr = os.path.join(td, "c.bin")
            run_gen(line, csr)
            dump = os.path.join(td, "d.bin")
            subprocess.run([os.path.join(REF, "div_dump"), csr, str(P), dump, "0", tol], check=True,
                           capture_output=True, env=dict(os.environ, OMP_NUM_THREADS=str(T)))
            parts = read_div(dump)
            stream = np.concatenate([mask_padding(b) for b, _ in parts]) if parts else np.zeros(0, np.uint8)
            key = "%s/T%d/P%d/tol%s" % (name, T, P, tol)
            out[key] = {"line": line, "T": T, "L": 4, "P": P, "tolerance": float(tol), "bytes": int(stream.size),
                        "stream": digest(stream)}
            print(key, stream.size, flush=True)
            if name == "test_txt":
                np.savez_compressed(os.path.join(HERE, "tiny_div_lossy.npz"), stream=stream)
    with open(os.path.join(HERE, "div_lossy.json"), "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)


def div_reorder_fixtures(gen):
    if not os.path.exists(os.path.join(REF, "div_dump")):
        sys.exit("oracle/_ref/div_dump missing: run `make -C oracle ref` first")
    cases = [("test_txt", 1, 97), ("test_txt", 4, 512), ("cfg1", 8, 4096), ("neigh_two", 2, 2048),
             ("batch_03", 5, 1000), ("tall", 4, 16384), ("simple_placement", 2, 300), ("stream_01", 3, 4096)]
    out = {}
    with tempfile.TemporaryDirectory() as td:
        for name, T, P in cases:
            line = gen[name]["line"]
            csr = os.path.join(td, "c.bin")
            run_gen(line, csr)
            dump = os.path.join(td, "d.bin")
            subprocess.run([os.path.join(REF, "div_dump"), csr, str(P), dump, "1"], check=True,
                           capture_output=True, env=dict(os.environ, OMP_NUM_THREADS=str(T)))
            seq = []
            for b, pos in read_div(dump):
                seq.append([int(np.frombuffer(b[q + 12:q + 16].tobytes(), dtype="<u4")[0]) for q in pos[:-1]])
            key = "%s/T%d/P%d" % (name, T, P)
            out[key] = {"line": line, "T": T, "P": P, "group_first_rows": seq}
            print(key, sum(len(s) for s in seq), flush=True)
    with open(os.path.join(HERE, "div_reorder.json"), "w") as f:
        json.dump(out, f, sort_keys=True)


def sell_fixtures(gen):
    if not os.path.exists(os.path.join(REF, "sell_dump")):
        sys.exit("oracle/_ref/sell_dump missing: run `make -C oracle ref` first")
    cases = [("test_txt", 1), ("test_txt", 3), ("test_txt", 8), ("cfg1", 4), ("neigh_two", 2),
             ("batch_03", 5), ("dense_rows", 1), ("tall", 4), ("stream_01", 7)]
    out = {}
    with tempfile.TemporaryDirectory() as td:
        for name, T in cases:
            line = gen[name]["line"]
            csr = os.path.join(td, "c.bin")
            run_gen(line, csr)
            dump = os.path.join(td, "s.bin")
            subprocess.run([os.path.join(REF, "sell_dump"), csr, dump], check=True, capture_output=True,
                           env=dict(os.environ, OMP_NUM_THREADS=str(T)))
            m = gen[name]["m"]
            with open(dump, "rb") as f:
                VL, TT, ncl, ne = struct.unpack("<4q", f.read(32))
                cptr = np.frombuffer(f.read(4 * (ncl + 1)), dtype=np.int32)
                ja = np.frombuffer(f.read(4 * ne), dtype=np.int32)
                a = np.frombuffer(f.read(8 * ne), dtype=np.float64)
                y = np.frombuffer(f.read(8 * m), dtype=np.float64)
            key = "%s/T%d" % (name, T)
            out[key] = {"line": line, "VL": int(VL), "T": T, "nnz_ext": int(ne),
                        "cluster_ptr": digest(cptr.astype(np.int64)), "ja": digest(ja.astype(np.int64)),
                        "a": digest(a), "y": digest(y)}
            print(key, int(ne), flush=True)
    with open(os.path.join(HERE, "sell.json"), "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)


def sell_sorted_fixtures(gen):
    if not os.path.exists(os.path.join(REF, "sell_sorted_dump")):
        sys.exit("oracle/_ref/sell_sorted_dump missing: run `make -C oracle ref` first")
    cases = [("test_txt", 1), ("test_txt", 3), ("test_txt", 8), ("cfg1", 4), ("neigh_two", 2),
             ("batch_03", 5), ("dense_rows", 1), ("tall", 4), ("stream_01", 7)]
    out = {}
    with tempfile.TemporaryDirectory() as td:
        for name, T in cases:
            line = gen[name]["line"]
            csr = os.path.join(td, "c.bin")
            run_gen(line, csr)
            dump = os.path.join(td, "s.bin")
            subprocess.run([os.path.join(REF, "sell_sorted_dump"), csr, dump], check=True, capture_output=True,
                           env=dict(os.environ, OMP_NUM_THREADS=str(T)))
            m = gen[name]["m"]
            with open(dump, "rb") as f:
                VL, TT, ncl, ne = struct.unpack("<4q", f.read(32))
                rng = np.frombuffer(f.read(16 * TT), dtype=np.int64).reshape(TT, 2)
                cptr = np.frombuffer(f.read(4 * (ncl + 1)), dtype=np.int32)
                perm = np.frombuffer(f.read(4 * m), dtype=np.int32)
                ja = np.frombuffer(f.read(4 * ne), dtype=np.int32)
                a = np.frombuffer(f.read(8 * ne), dtype=np.float64)
                y = np.frombuffer(f.read(8 * m), dtype=np.float64)
            key = "%s/T%d" % (name, T)
            out[key] = {"line": line, "VL": int(VL), "T": T, "nnz_ext": int(ne),
                        "windows": [int(v) for v in rng[:, 0]] + [int(rng[-1, 1])],
                        "perm": digest(perm), "cluster_ptr": digest(cptr.astype(np.int64)),
                        "ja": digest(ja.astype(np.int64)), "a": digest(a), "y": digest(y)}
            print(key, int(ne), flush=True)
    with open(os.path.join(HERE, "sell_sorted.json"), "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    if sys.argv[1:] in (["sell"], ["sell_sorted"], ["div_reorder"], ["div_lossy"]):
        with open(os.path.join(HERE, "generator.json")) as f:
            globals()[sys.argv[1] + "_fixtures"](json.load(f))
    else:
        main()
