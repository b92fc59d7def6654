"""DIV-RF encoding (host half of csr_to_format) against the reference's own
packet streams (golden fixtures written by tests/golden/make_fixtures.py from
the compiled reference, 4 value lanes = its AVX2 build), and the lossless
round trip that the reference checks on every conversion
(src/spmv_kernels/DIV/div.cpp:515-602): decoding every packet gives back the
CSR matrix exactly."""
import hashlib
import json
import os

import numpy as np
import pytest

from spmvdare import gen, hip
import pyoracle

HERE = os.path.dirname(os.path.abspath(__file__))
DIV = json.load(open(os.path.join(HERE, "golden", "div.json")))
GEN = json.load(open(os.path.join(HERE, "golden", "generator.json")))
REORDER = json.load(open(os.path.join(HERE, "golden", "div_reorder.json")))
LOSSY = json.load(open(os.path.join(HERE, "golden", "div_lossy.json")))


def digest(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()[:32]


@pytest.mark.parametrize("key", sorted(DIV))
def test_encoder_matches_reference_stream(key):
    f = DIV[key]
    a = gen.generate(gen.GenParams.parse(f["line"]), stats=False, map_values=False)
    b, off, grp = hip.div_encode(a.row_ptr, a.col_ind, a.values, a.n, T=f["T"], L=f["L"], P=f["P"])
    assert b.size == f["bytes"]
    assert digest(b) == f["stream"]  # the encoder writes the header padding as 0
    assert grp.size - 1 == sum(f["groups_per_partition"])


# ------------------------------------------------------------ lossy values

def test_lossy_oracle_matches_reference_decode(golden_dir):
    """The reference's own lossy stream (test_txt, T=4, P=512, tolerance
    1e-6) decodes to the oracle's rounding of every value."""
    a = gen.generate(gen.GenParams.parse(GEN["test_txt"]["line"]), stats=False, map_values=False)
    stream = np.load(os.path.join(golden_dir, "tiny_div_lossy.npz"))["stream"]
    rows, cols, vals = pyoracle.div_decode(stream, 4, a.nnz)
    assert rows.size == a.nnz
    # every (row, col) is unique, so it names its CSR entry
    csr_index = {(int(r), int(c)): j for r in range(a.m) for j, c in
                 zip(range(a.row_ptr[r], a.row_ptr[r + 1]), a.col_ind[a.row_ptr[r]:a.row_ptr[r + 1]])}
    order = np.array([csr_index[(int(r), int(c))] for r, c in zip(rows, cols)])
    expect = pyoracle.div_reduce_precision(a.values[order], 1e-6)
    np.testing.assert_array_equal(vals, expect)
    assert (vals != a.values[order]).any()


@pytest.mark.parametrize("key", sorted(LOSSY))
def test_lossy_encoder_matches_reference_stream(key):
    """The product's rounding equals the oracle's bit for bit, and encoding
    the rounded values gives the reference's lossy stream byte for byte."""
    f = LOSSY[key]
    a = gen.generate(gen.GenParams.parse(f["line"]), stats=False, map_values=False)
    r = hip.div_reduce_precision(a.values, f["tolerance"])
    sample = np.random.default_rng(0).choice(a.nnz, size=min(a.nnz, 20000), replace=False)
    np.testing.assert_array_equal(r[sample], pyoracle.div_reduce_precision(a.values[sample], f["tolerance"]))
    b, off, grp = hip.div_encode(a.row_ptr, a.col_ind, r, a.n, T=f["T"], L=f["L"], P=f["P"])
    assert b.size == f["bytes"]
    assert digest(b) == f["stream"]


@pytest.mark.parametrize("tol", [0.3, 1e-2, 1e-6, 2.0 ** -30, 1e-12, 1e-15])
def test_lossy_rounding_bounds(tol):
    """Error bound and bit pattern of the rounding on values of every
    magnitude and sign: |r - v| <= tol |v|, the low 52 - b fraction bits of r
    are zero (b = ceil(|log2 tol|)), and r is one of the two neighbours of v
    on that grid."""
    rng = np.random.default_rng(7)
    v = np.concatenate([rng.standard_normal(5000) * 10.0 ** rng.integers(-300, 300, 5000),
                        [0.0, -0.0, 1.0, -2.5, 0.1, 1 / 3, np.finfo(float).tiny, np.finfo(float).max / 4]])
    r = hip.div_reduce_precision(v, tol)
    np.testing.assert_array_equal(r, pyoracle.div_reduce_precision(v, tol))
    nz = v != 0
    assert (np.abs(r[nz] - v[nz]) <= tol * np.abs(v[nz])).all()
    keep = int(np.ceil(abs(np.log2(tol))))
    u, ru = v.view(np.uint64), r.view(np.uint64)
    if keep < 52:
        low = np.uint64((1 << (52 - keep)) - 1)
        assert ((ru & low) == 0).all()
        dn, up = u & ~low, (u + low) & ~low
        assert ((ru == dn) | (ru == up) | (u == 0) | (v == 0)).all()
    else:
        np.testing.assert_array_equal(r, v)
    np.testing.assert_array_equal(hip.div_reduce_precision(v, 0.0), v)


def test_lossy_tolerance_rejected():
    with pytest.raises(hip.SpmvHipError, match="tolerance"):
        hip.div_reduce_precision(np.ones(3), 1.0)


def median_key(cols):
    """The reference's median (lib/array_metrics.c:940-1000, method
    "median_unbiased" at q = 1/2): virtual index (N - 1) q + q / 3 + 1 / 3 =
    N / 2 into the sorted values, linear between neighbours, the maximum when
    the index reaches the last value -- half a position above numpy's
    median_unbiased, which puts it at N / 2 - 1/2. Truncated to an integer."""
    v = np.sort(np.asarray(cols, dtype=np.float64))
    N = v.size
    if N // 2 >= N - 1:
        return int(v[-1])
    if N % 2 == 0:
        return int(v[N // 2])
    return int(0.5 * v[N // 2] + 0.5 * v[N // 2 + 1])


@pytest.mark.parametrize("key", sorted(REORDER))
def test_packet_reordering_matches_reference(key):
    """The group order of DIV_PACKET_REORDERING=1 (div.cpp:377-395,455-498)
    against the reference's own reordered streams. Each group is named by the
    first row of its first packet. The key of the partition's k-th group is
    the median of the columns of the partition's k-th packet as median_key
    states it: the reference's sequence must be sorted by it, and the
    encoder's order must equal the reference's up to the order of equal keys
    (the reference sorts with an unstable quicksort; the encoder keeps the
    stream order among them)."""
    f = REORDER[key]
    a = gen.generate(gen.GenParams.parse(f["line"]), stats=False, map_values=False)
    b, off, grp = hip.div_encode(a.row_ptr, a.col_ind, a.values, a.n, T=f["T"], L=4, P=f["P"])
    order = hip.div_group_order(a.row_ptr, a.col_ind, a.values, a.n, T=f["T"], L=4, P=f["P"])
    ng = grp.size - 1
    assert sorted(order.tolist()) == list(range(ng))
    nv = np.array([int(np.frombuffer(b[o:o + 4].tobytes(), "<u4")[0]) for o in off[:-1]], dtype=np.int64)
    first_nnz = np.concatenate([[0], np.cumsum(nv)])
    row_min = np.array([int(np.frombuffer(b[off[p] + 12:off[p] + 16].tobytes(), "<u4")[0]) for p in grp[:-1]])
    g0 = 0
    for seq in f["group_first_rows"]:
        g1 = g0 + len(seq)
        keys = {}
        for k in range(g1 - g0):
            p = grp[g0] + k  # the partition's k-th packet
            cols = a.col_ind[first_nnz[p]:first_nnz[p + 1]]
            keys[int(row_min[g0 + k])] = median_key(cols)
        ref_keys = [keys[r] for r in seq]
        assert ref_keys == sorted(ref_keys), "the reference's order is not sorted by the median key"
        mine = [int(row_min[g]) for g in order[g0:g1]]
        assert [keys[r] for r in mine] == ref_keys
        natural = [int(r) for r in row_min[g0:g1]]
        assert mine == sorted(natural, key=lambda r: keys[r])  # stable: stream order among equal keys
        assert sorted(mine) == sorted(seq)
        g0 = g1
    assert g0 == ng


def test_tiny_stream_and_groups(golden_dir):
    ref = np.load(os.path.join(golden_dir, "tiny_div.npz"))
    a = gen.generate(gen.GenParams.parse(GEN["test_txt"]["line"]), stats=False, map_values=False)
    b, off, grp = hip.div_encode(a.row_ptr, a.col_ind, a.values, a.n, T=4, L=4, P=512)
    np.testing.assert_array_equal(b, ref["stream"])
    # packet-group starts as byte positions in the concatenated stream
    starts, base, t = [], 0, 0
    while "group_pos_%d" % t in ref:
        pos = ref["group_pos_%d" % t]
        starts += list(base + pos[:-1])
        base += int(pos[-1])
        t += 1
    assert list(off[grp[:-1]]) == starts


def roundtrip(rp, ci, va, n, T, L, P):
    b, off, grp = hip.div_encode(rp, ci, va, n, T=T, L=L, P=P)
    nnz = ci.size
    r, c, v = pyoracle.div_decode(b, L, nnz)
    assert r.size == nnz
    order = np.lexsort((c, r))
    rows = np.repeat(np.arange(rp.size - 1), np.diff(rp))
    np.testing.assert_array_equal(r[order], rows)
    np.testing.assert_array_equal(c[order], ci)
    np.testing.assert_array_equal(v[order], va)
    assert pyoracle.div_packet_sizes(b, L) == list(np.diff(off))


@pytest.mark.parametrize("name", ["test_txt", "batch_03", "tall", "dense_rows", "degree_clamp",
                                  "simple_placement", "run_sh_1"])
@pytest.mark.parametrize("T,L,P", [(1, 64, 16384), (4, 4, 512), (3, 7, 1000), (2, 1, 64), (16, 64, 4096)])
def test_roundtrip(name, T, L, P):
    a = gen.generate(gen.GenParams.parse(GEN[name]["line"]), stats=False)
    roundtrip(a.row_ptr, a.col_ind, a.values, a.n, T, L, P)


def test_roundtrip_repeated_values_and_empty_rows():
    """Replication factors > 1 (many equal values) and empty rows between
    packets, which the packetiser skips (div.cpp:356-357)."""
    rng = np.random.default_rng(1)
    m, n = 700, 900
    lens = rng.integers(0, 40, m)
    lens[rng.random(m) < 0.3] = 0
    rp = np.zeros(m + 1, np.int32)
    rp[1:] = np.cumsum(lens)
    ci = np.concatenate([np.sort(rng.choice(n, k, replace=False)) for k in lens]).astype(np.int32)
    va = rng.choice([0.5, -1.25, 3.0, 2.0 ** -30, 7.0], size=ci.size)
    va[::7] = rng.uniform(-1, 1, va[::7].size)
    for T, L, P in [(1, 4, 16384), (3, 64, 200), (5, 13, 777)]:
        roundtrip(rp, ci, va, n, T, L, P)


def _u32(b, at):
    return int(np.frombuffer(b[at:at + 4].tobytes(), dtype=np.uint32)[0])


def _le(b, at, nbytes):
    return int.from_bytes(b[at:at + nbytes].tobytes(), "little")


def _streams(pk, at, lane_sizes, L, il):
    """The L lane streams of a device packet's stream section (from byte
    `at`), one bytes object per lane: contiguous, or de-interleaved from rows
    of L dwords."""
    out, base = [], at
    for l in range(L):
        sz = int(lane_sizes[l])
        if il:
            idx = at + (np.arange(sz) >> 2) * 4 * L + 4 * l + (np.arange(sz) & 3)
            out.append(pk[idx].tobytes())
        else:
            out.append(pk[base:base + sz].tobytes())
            base += sz
    return out


@pytest.mark.parametrize("layout", [1, 2, 3, 4])
@pytest.mark.parametrize("name,T,L,P", [("test_txt", 4, 4, 512), ("cfg1", 2, 64, 1024), ("batch_03", 3, 32, 300),
                                        ("dense_rows", 1, 64, 16384), ("tall", 8, 64, 97)])
def test_device_layout_decodes_to_reference_packets(layout, name, T, L, P):
    """Every device layout of every packet against the reference packet:
    x slots (layouts 2, 4) are the packet's distinct columns in ascending
    order (runs maximal, one start bit per run) and every coordinate maps back
    to the (row, column) the reference packet holds at the same position;
    interleaved streams (3, 4) de-interleave to the reference lane streams;
    lengths and everything else are the reference's bytes. The decode below
    is restated here, not taken from the library."""
    slots, il = layout in (2, 4), layout in (3, 4)
    a = gen.generate(gen.GenParams.parse(GEN[name]["line"]), stats=False)
    b, off, grp = hip.div_encode(a.row_ptr, a.col_ind, a.values, a.n, T=T, L=L, P=P)
    r, c, v = pyoracle.div_decode(b, L, a.nnz)
    sb, soff = hip.div_device_encode(a.row_ptr, a.col_ind, a.values, a.n, T=T, L=L, P=P, layout=layout)
    assert soff.size == off.size and (soff % 16 == 0).all()
    k0 = 0
    for p in range(off.size - 1):
        ref = b[off[p]:off[p + 1]]
        pk = sb[soff[p]:soff[p + 1]]
        nv, nu, nrfs = _u32(ref, 0), _u32(ref, 4), _u32(ref, 24)
        ref_cb = (int(ref[20]) + int(ref[21]) + 7) // 8
        ref_lens_at = 28 + 4 * L + 8 * nrfs + nv * ref_cb
        lane_sizes = np.frombuffer(ref[28:28 + 4 * L].tobytes(), dtype=np.uint32)
        ref_streams = _streams(ref, ref_lens_at + nu, lane_sizes, L, False)
        if not slots:
            assert int(pk[22]) == int(il)
            np.testing.assert_array_equal(np.delete(pk[:ref_lens_at + nu], 22), np.delete(ref[:ref_lens_at + nu], 22))
            assert _streams(pk, (ref_lens_at + nu + 3) // 4 * 4 if il else ref_lens_at + nu, lane_sizes, L,
                            il) == ref_streams
            k0 += nv
            continue
        nv2, nu2, nrows, row_min, nslots = (_u32(pk, 4 * i) for i in range(5))
        rb, sbits, cb, il_flag = int(pk[20]), int(pk[21]), int(pk[22]), int(pk[23])
        nrfs2, nruns = _u32(pk, 24), _u32(pk, 28)
        assert (nv2, nu2, nrows, row_min, nrfs2, il_flag) == tuple(_u32(ref, 4 * i) for i in range(4)) + (nrfs,
                                                                                                              int(il))
        np.testing.assert_array_equal(pk[32:32 + 4 * L], ref[28:28 + 4 * L])            # lane sizes
        np.testing.assert_array_equal(pk[32 + 4 * L:32 + 4 * L + 8 * nrfs],
                                      ref[28 + 4 * L:28 + 4 * L + 8 * nrfs])                # rfs, unique per rf
        runs_at = 32 + 4 * L + 8 * nrfs
        bits_at = (runs_at + 4 * nruns + 7) // 8 * 8
        coords_at = bits_at + 8 * ((nslots + 63) // 64)
        run_col = np.frombuffer(pk[runs_at:runs_at + 4 * nruns].tobytes(), dtype=np.uint32)
        words = np.frombuffer(pk[bits_at:coords_at].tobytes(), dtype=np.uint64)
        starts = [s for s in range(nslots) if (int(words[s // 64]) >> (s % 64)) & 1]
        assert starts[0] == 0 and len(starts) == nruns
        slot_col = np.empty(nslots, dtype=np.int64)
        bounds = starts + [nslots]
        for i in range(nruns):
            slot_col[bounds[i]:bounds[i + 1]] = int(run_col[i]) + np.arange(bounds[i + 1] - bounds[i])
        cols = np.asarray(c[k0:k0 + nv], dtype=np.int64)
        np.testing.assert_array_equal(slot_col, np.unique(cols))                           # slots = distinct columns
        assert all(run_col[i] > slot_col[bounds[i] - 1] + 1 for i in range(1, nruns))     # runs are maximal
        assert sbits == (int(nslots - 1).bit_length() if nslots > 1 else 0)
        assert cb == max(1, (rb + sbits + 7) // 8) and rb == (int(nrows - 1).bit_length() if nrows > 1 else 0)
        for k in range(nv):
            cc = _le(pk, coords_at + k * cb, cb)
            assert row_min + (cc & ((1 << rb) - 1)) == r[k0 + k]
            assert slot_col[cc >> rb] == c[k0 + k]
        lens_at = coords_at + nv * cb
        np.testing.assert_array_equal(pk[lens_at:lens_at + nu], ref[ref_lens_at:ref_lens_at + nu])
        assert _streams(pk, (lens_at + nu + 3) // 4 * 4 if il else lens_at + nu, lane_sizes, L, il) == ref_streams
        k0 += nv
    assert k0 == a.nnz
