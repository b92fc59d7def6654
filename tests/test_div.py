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
