"""Pins the CPU oracle before it is trusted as the checker.

Closed forms (paths, stars, odd cycles, complete graphs, unit grids incl. the
uint64 overflow boundary), networkx's independent weighted Brandes, the RQ
pair formula against Brandes, and planted mistakes that must be caught.
"""
from math import comb

import numpy as np
import pytest

import oracle
from gpu_betweennesscentrality_amd import graphgen as gg


def csr_from_edges(n, edges, w=None):
    edges = np.asarray(edges, np.int64).reshape(-1, 2)
    w = np.ones(len(edges), np.int64) if w is None else np.asarray(w, np.int64)
    u = np.concatenate([edges[:, 0], edges[:, 1]])
    v = np.concatenate([edges[:, 1], edges[:, 0]])
    ww = np.concatenate([w, w])
    return oracle.CSR(n, u, v, ww)


def test_path_closed_form():
    n = 9
    bc, _ = oracle.brandes(csr_from_edges(n, [(i, i + 1) for i in range(n - 1)]))
    assert np.allclose(bc, [i * (n - 1 - i) for i in range(n)])


def test_star_closed_form():
    n = 11
    bc, _ = oracle.brandes(csr_from_edges(n, [(0, i) for i in range(1, n)]))
    assert bc[0] == (n - 1) * (n - 2) / 2 and np.all(bc[1:] == 0)


@pytest.mark.parametrize("n", [5, 9, 13])
def test_odd_cycle_closed_form(n):
    bc, _ = oracle.brandes(csr_from_edges(n, [(i, (i + 1) % n) for i in range(n)]))
    assert np.allclose(bc, (n - 1) * (n - 3) / 8)


def test_even_cycle_and_complete():
    for n in (8, 20):  # even cycle: antipodal pairs split 1/2 over two paths
        bc, _ = oracle.brandes(csr_from_edges(n, [(i, (i + 1) % n) for i in range(n)]))
        assert np.allclose(bc, (n - 2) ** 2 / 8)
    k = 7
    bc, _ = oracle.brandes(csr_from_edges(k, [(i, j) for i in range(k) for j in range(i + 1, k)]))
    assert np.all(bc == 0)


def test_unit_grid_counts_and_overflow():
    for k in (5, 12, 34):
        g = gg.unit_grid(k)
        d, s, h, ovf = oracle.sssp(oracle.CSR.from_arcgraph(g), 0)
        assert int(s[k * k - 1]) == comb(2 * k - 2, k - 1)
        assert d[k * k - 1] == 2 * k - 2 and h[k * k - 1] == 2 * k - 2
        assert not ovf
    g = gg.unit_grid(35)
    _, s, _, ovf = oracle.sssp(oracle.CSR.from_arcgraph(g), 0)
    assert ovf and int(s[35 * 35 - 1]) == 2**64 - 1


def test_hop_is_longest_dag_path():
    # two shortest 0->3 paths: 0-3 (w=4) and 0-1-2-3 (w=1+1+2): hop = 3
    csr = csr_from_edges(4, [(0, 3), (0, 1), (1, 2), (2, 3)], [4, 1, 1, 2])
    d, s, h, _ = oracle.sssp(csr, 0)
    assert d[3] == 4 and s[3] == 2 and h[3] == 3


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_networkx_weighted_brandes(seed):
    nx = pytest.importorskip("networkx")
    g = gg.road_graph(300, 420, seed=seed)
    G = nx.Graph()
    for a, b, w in zip(g.u.tolist(), g.v.tolist(), g.w.tolist()):
        G.add_edge(a, b, weight=w)
    ref = nx.betweenness_centrality(G, weight="weight", normalized=False)
    bc, _ = oracle.brandes(oracle.CSR.from_arcgraph(g))
    exp = np.array([ref[i] for i in range(g.n)])
    assert np.allclose(bc, exp, rtol=1e-12, atol=1e-9)


def test_networkx_grid_with_ties():
    nx = pytest.importorskip("networkx")
    g = gg.grid_graph(k=9, seed=3, wmax=3)  # many equal-length paths
    G = nx.Graph()
    for a, b, w in zip(g.u.tolist(), g.v.tolist(), g.w.tolist()):
        G.add_edge(a, b, weight=w)
    ref = nx.betweenness_centrality(G, weight="weight", normalized=False)
    bc, _ = oracle.brandes(oracle.CSR.from_arcgraph(g))
    assert np.allclose(bc, [ref[i] for i in range(g.n)], rtol=1e-12, atol=1e-9)


def test_rq_pairs_equals_brandes():
    g = gg.grid_graph(k=8, seed=5, wmax=4)
    csr = oracle.CSR.from_arcgraph(g)
    pts = np.arange(0, g.n, 5, dtype=np.int32)
    bc, _ = oracle.brandes(csr)
    assert np.allclose(oracle.rq_pairs(csr, pts), bc[pts], rtol=1e-12, atol=1e-12)


def test_planted_mistakes_are_caught():
    """A Brandes without the '+1' term, or a count that takes the max instead
    of the sum, must disagree with the closed forms the checks use."""
    n = 9
    csr = csr_from_edges(n, [(i, (i + 1) % n) for i in range(n)])
    d = np.zeros(n, np.int64)
    # dependency without the (1 + delta) term: delta[u] = sum sigma_u/sigma_v * delta_v
    bc_wrong = np.zeros(n)
    for s in range(n):
        dd, sg, _, _ = oracle.sssp(csr, s)
        order = np.argsort(dd, kind="stable")[::-1]
        delta = np.zeros(n)
        for u in order:
            for a in range(csr.row[u], csr.row[u + 1]):
                v = csr.col[a]
                if dd[v] == dd[u] + csr.w[a]:
                    delta[u] += sg[u] / sg[v] * delta[v]
        bc_wrong += 0.5 * delta
    assert not np.allclose(bc_wrong, (n - 1) * (n - 3) / 8)
    # counting by max instead of sum on a unit grid
    k = 6
    g = gg.unit_grid(k)
    _, s, _, _ = oracle.sssp(oracle.CSR.from_arcgraph(g), 0)
    assert int(s[k * k - 1]) != 1 and int(s[k * k - 1]) == comb(2 * k - 2, k - 1)
    del d


def test_label_bc_on_reference_rows_equals_brandes():
    """The reference-method CPU path (label queries over the reference's own
    TreeDecomp rows, then layered Brandes) equals Dijkstra-Brandes; a row with
    one entry changed does not."""
    from pathlib import Path

    golden = Path(__file__).resolve().parent / "golden"
    full = dict(np.load(golden / "grid1k_tcfull.npz"))
    tc = np.load(golden / "grid1k_tc.npz")
    csr = oracle.CSR(int(tc["n"]), tc["u"], tc["v"], tc["w"])
    ref, _ = oracle.brandes(csr)
    got = oracle.label_bc(csr, full, threads=4)
    assert np.all(np.abs(got - ref) <= 1e-9 * np.maximum(np.abs(ref), 1.0))
    bad = dict(full, dis=full["dis"].copy())
    row = np.arange(full["label_off"][500], full["label_off"][501])
    k = row[np.flatnonzero(full["dis"][row] < 999_999_999)[0]]
    bad["dis"][k] //= 2  # a much shorter restricted path that does not exist
    wrong = oracle.label_bc(csr, bad, threads=4)
    assert not np.all(np.abs(wrong - ref) <= 1e-9 * np.maximum(np.abs(ref), 1.0))


def test_label_bc_on_built_rows_equals_brandes():
    import gpu_betweennesscentrality_amd as bcx

    g = gg.road_graph(3000, 3750, seed=21)
    G = bcx.Graph.from_arcs(g.n, g.u, g.v, g.w)
    tree = bcx.Index.build_mhl(G).export()
    csr = oracle.CSR.from_arcgraph(g)
    src = gg.permutation(3, 1, g.n)[:300].astype(np.int32)
    ref, _ = oracle.brandes(csr, src)
    got = oracle.label_bc(csr, tree, src, threads=3)
    assert np.all(np.abs(got - ref) <= 1e-9 * np.maximum(np.abs(ref), 1.0))
