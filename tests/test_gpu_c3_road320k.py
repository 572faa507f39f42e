"""Parity at the 320,000-vertex road-graph configuration, the headline
workload (batched MHL-QBC, MHLQBC.cu:495-537), and RQ serving on it
(RQ.cu:465-531): sampled sources against the oracle (every wide unpack level
runs), one headline launch over every source (5,000 tiles drawing on 512
state slots) in the two-workgroups-per-unit tile-kernel form against the
one-per-unit form (bitwise) and against the label scan, the forward pass's
rows bit for bit, and an RQ request alone and co-batched."""
import numpy as np
import pytest

import gpu_betweennesscentrality_amd as bcx
from gpu_betweennesscentrality_amd import graphgen as gg
import oracle
from _bigcfg import close_bc, max_rel, oracle_rows, rows_equal, sample_sources, widest_level, workload

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def cfg():
    g, csr, G, ix, ctx = workload("road320k")
    yield g, csr, ix, ctx
    ctx.close()


def test_has_wide_unpack_levels(cfg):
    g, csr, ix, ctx = cfg
    assert widest_level(ix) >= ctx.tuning()["unpack_wide"]


def test_sampled_sources_vs_oracle(cfg):
    g, csr, ix, ctx = cfg
    src = sample_sources(ix, g.n, 64, 64, seed=43)
    want, _ = oracle.brandes(csr, src, threads=oracle.default_threads())
    got, ovf = ctx.bc_sources(src, bcx.MODE_SWEEP)
    assert not ovf
    assert close_bc(got, want), f"max rel err {max_rel(got, want):.3e}"


def test_full_launch_two_forms_bitwise_and_label_scan(cfg):
    g, csr, ix, ctx = cfg
    batch = ctx.tuning()["sweep_batch"]
    _, cus = ctx.sweep_slots()
    assert batch // 64 > 2 * cus, "the headline launch holds more tiles than state slots"
    src = ix.tree_order()[:batch].astype(np.int32)
    ctx.reset_timing()
    two, o1 = ctx.bc_sources(src, bcx.MODE_SWEEP)
    assert ctx.timing()["scan_launches"] == 1
    ctx.set_tuning(sweep_form=bcx.SWEEP_FORM_ONE_PER_CU)
    one, o2 = ctx.bc_sources(src, bcx.MODE_SWEEP)
    ctx.set_tuning(sweep_form=bcx.SWEEP_FORM_AUTO)
    assert not o1 and not o2
    assert np.array_equal(two.view(np.uint64), one.view(np.uint64))
    scan, _ = ctx.bc_sources(src, bcx.MODE_BATCHED)
    assert close_bc(two, scan), f"max rel err {max_rel(two, scan):.3e}"


def test_rows_bit_exact_both_forms(cfg):
    g, csr, ix, ctx = cfg
    src = sample_sources(ix, g.n, 64, 8, seed=44)
    want = oracle_rows(csr, src)
    for form in (bcx.SWEEP_FORM_ONE_PER_CU, bcx.SWEEP_FORM_TWO_PER_CU):
        ctx.set_tuning(sweep_form=form)
        err = rows_equal(ctx.query_rows(src, bcx.MODE_SWEEP), want)
        assert err is None, f"form {form}: {err}"
    ctx.set_tuning(sweep_form=bcx.SWEEP_FORM_AUTO)
    assert rows_equal(ctx.query_rows(src[:64], bcx.MODE_BATCHED), want[:64]) is None


def test_rq_alone_and_co_batched(cfg):
    """A 5,000-vertex uniform request alone and joining a 20,000-vertex
    BFS-ball request in the middle of its cycle: bitwise the same, and both
    equal to the full BC at their vertices."""
    g, csr, ix, ctx = cfg
    row = np.zeros(g.n + 1, np.int64)
    np.add.at(row, g.u.astype(np.int64) + 1, 1)
    row = np.cumsum(row)
    a = np.sort(gg.permutation(45, 1, g.n)[:5000]).astype(np.int32)
    centre = 1234
    seen, frontier, order = {centre}, [centre], [centre]
    while frontier and len(order) < 20000:
        nxt = []
        for x in frontier:
            for y in g.v[row[x]:row[x + 1]].tolist():
                if y not in seen:
                    seen.add(y)
                    nxt.append(y)
                    order.append(y)
        frontier = nxt
    b = np.sort(np.array(order[:20000], np.int32))

    def run(srv, rid):
        while True:
            srv.step()
            r = srv.poll(rid)
            if r is not None:
                return r

    srv = bcx.RQServer(ctx, 1 << 34)
    alone = run(srv, srv.submit(a))
    srv.close()
    srv = bcx.RQServer(ctx, 1 << 34)
    rb = srv.submit(b)
    for _ in range(3):
        srv.step()
    ra = srv.submit(a)
    res = {}
    while len(res) < 2:
        srv.step()
        for rid in (ra, rb):
            if rid not in res:
                r = srv.poll(rid)
                if r is not None:
                    res[rid] = r
    srv.close()
    assert np.array_equal(alone.view(np.uint64), res[ra].view(np.uint64))
    full, ovf = ctx.full_bc(bcx.MODE_SWEEP)
    assert not ovf
    assert close_bc(alone, full[a]), f"max rel err {max_rel(alone, full[a]):.3e}"
    assert close_bc(res[rb], full[b]), f"max rel err {max_rel(res[rb], full[b]):.3e}"
