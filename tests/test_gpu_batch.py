"""Batched SpMV: one persistent, ticket-scheduled launch over many
generated matrices of uneven sizes, directly and replayed from a hipGraph.
Every member's y is checked against the oracle (normwise <= 1e-13) and, for
SELL members, bit-identical to the member's own standalone launch. DIV-RF
members run in the sequence's second persistent kernel."""
import json
import os

import numpy as np
import pytest

from spmvdare import gen, hip
import pyoracle

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
GEN = json.load(open(os.path.join(HERE, "golden", "generator.json")))
VARIANTS = [("csr_vec", {}), ("sell_c_s", {"sigma": 256}), ("sell", {}), ("sell_c_s", {"sigma": 4096}),
            ("csr_vec", {"block_nnz": 128}), ("csr_vec", {"block_nnz": 256, "lanes_per_row": 4}),
            ("div_rf", {"div_packet_vals": 1024, "unroll": 8}),
            ("div_rf", {"div_packet_vals": 512, "div_layout": 1, "div_lanes": 32, "unroll": 4})]


@pytest.fixture(scope="module")
def members():
    hip.init(0)
    names = sorted(k for k in GEN if k.startswith("batch_"))[:16] + ["test_txt", "dense_rows", "run_sh_1"]
    out = []
    for i, name in enumerate(names):
        a = gen.generate(gen.GenParams.parse(GEN[name]["line"]), stats=False)
        x = gen.make_x(a.n, 1000 + i)
        kind, kw = VARIANTS[i % len(VARIANTS)]
        f = hip.csr_to_format(a.row_ptr, a.col_ind, a.values, a.m, a.n, a.nnz, kind=kind, **kw)
        dx = hip.DeviceArray.from_host(x)
        dy = hip.DeviceArray(a.m)
        y_ref = pyoracle.csr_naive(a.row_ptr, a.col_ind, a.values, x)
        y_alone = np.empty(a.m)
        f.spmv(x, y_alone)
        out.append(dict(name=name, kind=kind, fmt=f, dx=dx, dy=dy, y_ref=y_ref, y_alone=y_alone, m=a.m))
        print("member", i, name, kind, kw, a.m, a.nnz, flush=True)
    return out


def check(members):
    for mb in members:
        y = mb["dy"].download()
        assert pyoracle.normwise_rel_err(y, mb["y_ref"]) <= 1e-13, mb["name"]
        if mb["kind"] in ("sell", "sell_c_s"):
            np.testing.assert_array_equal(y, mb["y_alone"])


@pytest.mark.parametrize("use_graph", [False, True])
@pytest.mark.parametrize("tile", [1, 16, 64])
def test_batch_matches_members(members, use_graph, tile):
    b = hip.Batch(tile)
    for mb in members:
        b.add(mb["fmt"], mb["dx"], mb["dy"])
    b.finalize(use_graph=use_graph)
    info = b.info()
    print("batch", info, flush=True)
    assert info["matrices"] == len(members)
    for rep in range(3):
        for mb in members:
            mb["dy"].upload(np.full(mb["m"], np.nan))
        b.run()
        b.synchronize()
        check(members)
        # eager work between graph replays
        members[0]["fmt"].spmv_device(members[0]["dx"], members[0]["dy"])
        hip.synchronize()
    ms = b.time(2, 5)
    ms_pm = b.time(2, 5, per_matrix=True)
    print("batched %.3f ms, per-matrix launches %.3f ms" % (np.median(ms), np.median(ms_pm)), flush=True)
    check(members)
    b.destroy()


def test_batch_with_block512_member(members):
    """A member with block_nnz = 512 switches the batch to the kernel with the
    larger LDS staging; every member must still be right."""
    a = gen.generate(gen.GenParams.parse(GEN["cfg1"]["line"]), stats=False)
    x = gen.make_x(a.n, 77)
    f = hip.csr_to_format(a.row_ptr, a.col_ind, a.values, a.m, a.n, a.nnz, kind="csr_vec", block_nnz=512)
    big = dict(name="cfg1_b512", kind="csr_vec", fmt=f, dx=hip.DeviceArray.from_host(x), dy=hip.DeviceArray(a.m),
               y_ref=pyoracle.csr_naive(a.row_ptr, a.col_ind, a.values, x), m=a.m)
    group = members[:7] + [big]
    b = hip.Batch(4)
    for mb in group:
        b.add(mb["fmt"], mb["dx"], mb["dy"])
    b.finalize(use_graph=True)
    for rep in range(2):
        for mb in group:
            mb["dy"].upload(np.full(mb["m"], np.nan))
        b.run()
        b.synchronize()
        check(group)
    b.destroy()


@pytest.mark.parametrize("use_graph,queues", [(False, 64), (True, 8), (True, 1), (True, 1024)])
def test_batch_split_rows(members, use_graph, queues):
    """Members whose long rows are split into segments (split_nnz 256): the
    segments run as work items of the persistent kernel and the finalize
    pass adds each row's partials; y must match the oracle and the member's
    own standalone launch (the same segments, the same order)."""
    group = []
    for i, name in enumerate(["dense_rows", "batch_03", "batch_07", "run_sh_1"]):
        a = gen.generate(gen.GenParams.parse(GEN[name]["line"]), stats=False)
        x = gen.make_x(a.n, 500 + i)
        f = hip.csr_to_format(a.row_ptr, a.col_ind, a.values, a.m, a.n, a.nnz, kind="csr_vec", block_nnz=128,
                              split_nnz=256)
        y_alone = np.empty(a.m)
        f.spmv(x, y_alone)
        group.append(dict(name=name, kind="csr_vec_split", fmt=f, dx=hip.DeviceArray.from_host(x),
                          dy=hip.DeviceArray(a.m), y_ref=pyoracle.csr_naive(a.row_ptr, a.col_ind, a.values, x),
                          y_alone=y_alone, m=a.m, nsplit=f.info()["num_split_rows"]))
    assert sum(g["nsplit"] for g in group) > 0
    b = hip.Batch(0)
    b.set_queues(queues)
    for mb in group:
        b.add(mb["fmt"], mb["dx"], mb["dy"])
    b.finalize(use_graph=use_graph)
    for rep in range(3):
        for mb in group:
            mb["dy"].upload(np.full(mb["m"], np.nan))
        b.run()
        b.synchronize()
        for mb in group:
            y = mb["dy"].download()
            assert pyoracle.normwise_rel_err(y, mb["y_ref"]) <= 1e-13, mb["name"]
            np.testing.assert_array_equal(y, mb["y_alone"])
    b.destroy()


@pytest.mark.parametrize("wgs", [1, 3, 8])
def test_batch_grid_sizes(members, wgs):
    """Any resident grid (workgroups per CU) computes the same y: the tickets
    cover every unit exactly once."""
    b = hip.Batch(2)
    b.set_grid(wgs)
    for mb in members:
        b.add(mb["fmt"], mb["dx"], mb["dy"])
    b.finalize(use_graph=True)
    print("batch", b.info(), flush=True)
    for mb in members:
        mb["dy"].upload(np.full(mb["m"], np.nan))
    b.run()
    b.synchronize()
    check(members)
    b.destroy()
    bad = hip.Batch()
    with pytest.raises(hip.SpmvHipError):
        bad.set_grid(9)
    for nq in (0, 3, 2048):
        with pytest.raises(hip.SpmvHipError):
            bad.set_queues(nq)
    bad.destroy()


@pytest.mark.parametrize("queues,tile", [(256, 1), (8, 3), (1024, 1)])
def test_batch_xcd_ranges(members, queues, tile):
    """With one contiguous unit range per XCD every unit is still computed
    exactly once (queues / 8 counters per range)."""
    b = hip.Batch(tile)
    b.set_queues(queues)
    b.set_xcd_ranges(True)
    for mb in members:
        b.add(mb["fmt"], mb["dx"], mb["dy"])
    b.finalize(use_graph=True)
    for rep in range(2):
        for mb in members:
            mb["dy"].upload(np.full(mb["m"], np.nan))
        b.run()
        b.synchronize()
        check(members)
    b.destroy()


def test_batch_rejects_unbatchable(members):
    a = gen.generate(gen.GenParams.parse(GEN["test_txt"]["line"]), stats=False)
    b = hip.Batch()
    for kind, kw in (("csr_naive", {}), ("csr_vec", {"block_nnz": 1024})):
        f = hip.csr_to_format(a.row_ptr, a.col_ind, a.values, a.m, a.n, a.nnz, kind=kind, **kw)
        dx = hip.DeviceArray(a.n)
        dy = hip.DeviceArray(a.m)
        with pytest.raises(hip.SpmvHipError):
            b.add(f, dx, dy)


@pytest.mark.parametrize("use_graph", [False, True])
def test_batch_div_members_only(members, use_graph):
    """A batch of DIV-RF members only (the CSR / SELL kernel has no units)."""
    sel = [mb for mb in members if mb["kind"] == "div_rf"]
    assert len(sel) >= 2
    b = hip.Batch()
    for mb in sel:
        b.add(mb["fmt"], mb["dx"], mb["dy"])
    b.finalize(use_graph=use_graph)
    for rep in range(3):
        for mb in sel:
            mb["dy"].upload(np.full(mb["m"], np.nan))
        b.run()
        b.synchronize()
        check(sel)
    b.destroy()


def test_batch_rejects_slot_layout_div():
    a = gen.generate(gen.GenParams.parse(GEN["test_txt"]["line"]), stats=False)
    f = hip.csr_to_format(a.row_ptr, a.col_ind, a.values, a.m, a.n, a.nnz, kind="div_rf", div_layout=2)
    dx = hip.DeviceArray.from_host(gen.make_x(a.n, 5))
    dy = hip.DeviceArray(a.m)
    b = hip.Batch()
    with pytest.raises(hip.SpmvHipError):
        b.add(f, dx, dy)
    b.destroy()


@pytest.mark.parametrize("ranges", [False, True])
def test_batch_empty_blocks_at_odd_offsets(members, ranges):
    """Empty row blocks that start at an odd nonzero index (a row of 300
    entries, then an empty row, which starts a block of its own): the paired
    stream loads still run once for them, and the wave must read its next
    ticket exactly once per unit, or a unit of another member is never run.
    Thousands of such blocks among the regular members; every y checked."""
    from test_gpu_formats import csr_from_lengths
    rp, ci, va = csr_from_lengths([3, 300, 0] * 2000, 5000, 7)
    m = rp.size - 1
    f = hip.csr_to_format(rp, ci, va, m, 5000, ci.size, kind="csr_vec")
    x = gen.make_x(5000, 77)
    dx, dy = hip.DeviceArray.from_host(x), hip.DeviceArray(m)
    y_ref = pyoracle.csr_naive(rp, ci, va, x)
    b = hip.Batch(1)
    b.set_xcd_ranges(ranges)
    b.add(f, dx, dy)
    for mb in members:
        b.add(mb["fmt"], mb["dx"], mb["dy"])
    b.finalize(use_graph=True)
    for rep in range(3):
        dy.upload(np.full(m, np.nan))
        for mb in members:
            mb["dy"].upload(np.full(mb["m"], np.nan))
        b.run()
        b.synchronize()
        assert pyoracle.normwise_rel_err(dy.download(), y_ref) <= 1e-13
        check(members)
    b.destroy()
