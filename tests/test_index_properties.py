"""Property tests of the host index (no GPU): on random small connected
graphs with many equal-length paths and random contraction orders, every
pair query of the count index (TreeDecomp.cpp:953-1020) and of the H2H
index (QBC.cu:271-331 restated on the host) equals the oracle's Dijkstra
with counts, and the index image and both host-array seams give back an
index that answers the same."""
import numpy as np
from hypothesis import HealthCheck, given, settings, strategies as st

import gpu_betweennesscentrality_amd as bcx
from gpu_betweennesscentrality_amd import graphgen as gg
import oracle


@st.composite
def graphs(draw):
    n = draw(st.integers(2, 36))
    # A random spanning tree keeps the graph connected; extra edges add cycles.
    parent = [draw(st.integers(0, v - 1)) for v in range(1, n)]
    edges = {(min(v, p), max(v, p)) for v, p in zip(range(1, n), parent)}
    extra = draw(st.integers(0, 2 * n))
    for _ in range(extra):
        a, b = draw(st.integers(0, n - 1)), draw(st.integers(0, n - 1))
        if a != b:
            edges.add((min(a, b), max(a, b)))
    edges = sorted(edges)
    wmax = draw(st.sampled_from([1, 2, 3, 1000]))  # small ranges: many ties
    w = [draw(st.integers(1, wmax)) for _ in edges]
    order = draw(st.permutations(list(range(n))))
    u = np.array([a for a, b in edges] + [b for a, b in edges], np.int32)
    v = np.array([b for a, b in edges] + [a for a, b in edges], np.int32)
    ww = np.array(w + w, np.int32)
    return gg.ArcGraph(n, u, v, ww), np.array(order, np.int32)


def _check_all_pairs(ix, csr, n, h2h=False):
    for s in range(n):
        d, sg, hp, _ = oracle.sssp(csr, s)
        for t in range(n):
            if h2h:
                assert ix.h2h_query(s, t) == d[t], (s, t)
            else:
                assert ix.query(s, t)[:3] == (d[t], sg[t], hp[t]), (s, t)


@settings(max_examples=60, deadline=None, suppress_health_check=[HealthCheck.too_slow])
@given(graphs())
def test_random_graphs_all_pairs(case):
    g, order = case
    csr = oracle.CSR.from_arcgraph(g)
    G = bcx.Graph.from_arcs(g.n, g.u, g.v, g.w)
    mhl = bcx.Index.build_mhl(G)
    _check_all_pairs(mhl, csr, g.n)
    ordered = bcx.Index.build_order(G, order)
    _check_all_pairs(ordered, csr, g.n)
    ordered.build_h2h()
    _check_all_pairs(ordered, csr, g.n, h2h=True)


@settings(max_examples=30, deadline=None, suppress_health_check=[HealthCheck.too_slow])
@given(graphs())
def test_random_graphs_round_trips(tmp_path_factory, case):
    g, order = case
    G = bcx.Graph.from_arcs(g.n, g.u, g.v, g.w)
    ix = bcx.Index.build_order(G, order)
    ix.build_h2h()
    path = tmp_path_factory.mktemp("img") / "ix.bin"
    ix.save(path)
    loaded = bcx.Index.load(path)
    seam = bcx.Index.from_arrays(ix.export(), G)
    h = ix.h2h_export()
    h["parent"] = ix.tree()[0]
    target = bcx.Index.build_order(G, order)
    target.h2h_from_arrays(h)
    for s in range(g.n):
        for t in range(g.n):
            want = ix.query(s, t)
            assert loaded.query(s, t) == want and seam.query(s, t) == want, (s, t)
            assert target.h2h_query(s, t) == ix.h2h_query(s, t) == loaded.h2h_query(s, t), (s, t)
