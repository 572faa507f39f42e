"""The H2H host-array seam (bcx_index_h2h_from_arrays), the input side of
moveDisIndexToDevice (QBC.cu:11-127): QBC's distance index as the reference's
Graph holds it after makeIndex, optimizeVS and SVSP_P2H
(src/buildIndex.cpp:208-278, 313-386, 565-719).  The fixture
grid1k_h2hfull.npz is that structure dumped by the reference's own host code
(oracle/_ref/tc_ref h2hfull) for the seeded grid on its min-degree order."""
import numpy as np
import pytest

import gpu_betweennesscentrality_amd as bcx
import oracle
from gpu_betweennesscentrality_amd import graphgen as gg


@pytest.fixture(scope="module")
def grid(golden_dir):
    g = gg.make("grid1k")
    G = bcx.Graph.from_arcs(g.n, g.u, g.v, g.w)
    ref = dict(np.load(golden_dir / "grid1k_h2hfull.npz"))
    return g, G, ref


def _lists(a, n):
    """{branch vertex: {vertex: sorted positions}} of an exported P2H table."""
    out = {}
    for j, b in enumerate(a["branch"].tolist()):
        o = a["p2h_off"][j * n:(j + 1) * n + 1]
        out[b] = {s: sorted(a["p2h"][o[s]:o[s + 1]].tolist()) for s in range(n) if o[s + 1] > o[s]}
    return out


def test_reference_h2h_index_gives_dijkstra_distances(grid):
    g, G, ref = grid
    ix = bcx.Index.build_order(G, ref["order"])
    ix.h2h_from_arrays(ref)
    csr = oracle.CSR.from_arcgraph(g)
    for s in range(0, g.n, 5):
        d = oracle.sssp(csr, s)[0]
        got = np.array([ix.h2h_query(s, t) for t in range(g.n)], np.int64)
        assert np.array_equal(got, d), f"source {s}"


def test_own_h2h_build_equals_the_reference_structure(grid):
    """Built from the same order, bcx's H2H distance rows equal the
    reference's bit for bit, its separators are the reference's sets (the
    reference emits them in hash-map order), and the P2H tables hold the same
    lists for the branch vertices both keep (the reference draws its sample
    pairs from the clock, bcx from a fixed seed)."""
    g, G, ref = grid
    ix = bcx.Index.build_order(G, ref["order"])
    ix.build_h2h()
    own = ix.h2h_export()
    assert np.array_equal(own["gdis_off"], ref["gdis_off"]) and np.array_equal(own["gdis"], ref["gdis"])
    assert np.array_equal(own["sep_off"], ref["sep_off"])
    for v in range(g.n):
        a, b = own["sep_off"][v], own["sep_off"][v + 1]
        assert sorted(own["sep"][a:b].tolist()) == sorted(ref["sep"][a:b].tolist()), f"vertex {v}"
    mine, theirs = _lists(own, g.n), _lists(ref, g.n)
    both = set(mine) & set(theirs)
    assert both, "no branch vertex in common"
    for b in both:
        assert mine[b] == theirs[b], f"branch vertex {b}"


def test_round_trip_through_the_seam(grid):
    g, G, ref = grid
    src = bcx.Index.build_order(G, ref["order"])
    src.build_h2h()
    a = src.h2h_export()
    a["parent"] = src.tree()[0]
    dst = bcx.Index.build_order(G, ref["order"])
    dst.h2h_from_arrays(a)
    b = dst.h2h_export()
    for k in a:
        if k != "parent":
            assert np.array_equal(a[k], b[k]), k
    pairs = gg.permutation(9, 1, g.n)[:300].reshape(-1, 2).tolist()
    assert [src.h2h_query(s, t) for s, t in pairs] == [dst.h2h_query(s, t) for s, t in pairs]
    assert src.h2h_scan_stats(4096, 3) == dst.h2h_scan_stats(4096, 3)


@pytest.mark.parametrize("damage", ["parent", "row_length", "sep_range", "p2h_outside", "branch_twice",
                                    "negative_distance"])
def test_damaged_arrays_are_refused_and_leave_the_index(grid, damage):
    g, G, ref = grid
    ix = bcx.Index.build_order(G, ref["order"])
    ix.h2h_from_arrays(ref)
    before = ix.h2h_export()
    a = {k: np.array(v, copy=True) for k, v in ref.items()}
    n = g.n
    if damage == "parent":
        v = int(np.flatnonzero(a["parent"] >= 0)[0])
        a["parent"][v] = (a["parent"][v] + 1) % n
    elif damage == "row_length":
        a["gdis_off"][5] += 1  # row 4 one entry longer, row 5 one shorter
    elif damage == "sep_range":
        v = int(np.argmax(np.diff(a["gdis_off"])))
        a["sep"][a["sep_off"][v]] = int(np.diff(a["gdis_off"])[v]) + 1  # past v's own position
    elif damage == "p2h_outside":
        par = a["parent"]
        b = int(a["branch"][0])
        below = np.zeros(n, bool)
        for s in range(n):  # vertices below b
            x = s
            while x >= 0 and x != b:
                x = par[x]
            below[s] = x == b and s != b
        s = int(np.flatnonzero(~below & (np.arange(n) != b))[0])
        o = a["p2h_off"]
        ins = int(o[s + 1])
        a["p2h"] = np.insert(a["p2h"], ins, 0)
        o[s + 1:] += 1
    elif damage == "branch_twice":
        a["branch"] = np.array([a["branch"][0], a["branch"][0]], np.int32)
        o0 = a["p2h_off"][:n + 1]
        a["p2h_off"] = np.concatenate([o0[:-1], o0 + o0[-1]])
        a["p2h"] = np.concatenate([a["p2h"][:o0[-1]], a["p2h"][:o0[-1]]])
    elif damage == "negative_distance":
        a["gdis"][10] = -1
    with pytest.raises(bcx.BcxError):
        ix.h2h_from_arrays(a)
    after = ix.h2h_export()
    for k in before:
        assert np.array_equal(before[k], after[k]), k
