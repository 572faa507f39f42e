"""The allocation-free probe (bcx_query_dist_count_noalloc) against the
entry point it mirrors (bcx_query_dist_count), through the C ABI: the same
distance, count and hop rows bit for bit on two calls in a row, the second on
a smaller graph, and the overflow flag on the saturating unit grid."""
import numpy as np
import pytest

import gpu_betweennesscentrality_amd as bcx
from gpu_betweennesscentrality_amd import graphgen as gg

pytestmark = pytest.mark.gpu


def _ctx(g):
    G = bcx.Graph.from_arcs(g.n, g.u, g.v, g.w)
    ix = bcx.Index.build_mhl(G)
    return bcx.Context(G, ix, 0)


def _same(a, b):
    return all(np.array_equal(x, y) for x, y in zip(a[:3], b[:3])) and a[3] == b[3]


def test_noalloc_matches_query_on_two_calls_smaller_second():
    big = _ctx(gg.road_graph(20_000, 25_000, seed=11))
    small = _ctx(gg.make("grid1k"))
    for s in (17, 19_999):
        assert _same(big.query_noalloc(s), big.query(s)), f"road20k source {s}"
    for s in (0, 1023, 512):  # smaller graph after the larger: stale rows must not leak
        assert _same(small.query_noalloc(s), small.query(s)), f"grid1k source {s}"
    assert _same(big.query_noalloc(5), big.query(5))


def test_noalloc_overflow_flag_and_bad_source():
    ctx = _ctx(gg.unit_grid(35))
    got = ctx.query_noalloc(0)
    want = ctx.query(0)
    assert got[3] and _same(got, want)
    with pytest.raises(bcx.BcxError):
        ctx.query_noalloc(35 * 35)
