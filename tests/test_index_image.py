"""The persisted index image, its validation on load, the graph fingerprint,
the host-array seam (the reference's TreeDecomp handed over) and the
distance bound, all on the CPU (no device call)."""
import struct
from pathlib import Path

import numpy as np
import pytest

import gpu_betweennesscentrality_amd as bcx
from gpu_betweennesscentrality_amd import graphgen as gg

GOLDEN = Path(__file__).resolve().parent / "golden"
# Vectors of an image in order, with their element types (the H2H vectors
# follow when the header's H2H flag is set).
VECS = [("perm", "i4"), ("depth", "i4"), ("parent", "i4"), ("row_off", "i8"), ("dis", "i4"), ("cnt", "u8"),
        ("hop", "u4"), ("bag_off", "i8"), ("bag", "i4"), ("bag_len", "i4"), ("bag_cnt", "u8"),
        ("bag_hop", "u4"), ("gdis", "i4"), ("sep_off", "i8"), ("sep", "i4"), ("p2h_node", "i4"),
        ("p2h_base", "i8"), ("p2h_off", "i8"), ("p2h", "i4")]
HEAD = 8 + 16 + 16 + 8


def vector_offsets(raw: bytes) -> dict:
    """name -> (offset of the first element, element count, dtype)."""
    has_h2h = struct.unpack_from("<i", raw, 8 + 16 + 16)[0]
    out, o = {}, HEAD
    for name, dt in VECS[: 19 if has_h2h else 12]:
        k = struct.unpack_from("<q", raw, o)[0]
        out[name] = (o + 8, k, np.dtype(dt))
        o += 8 + k * np.dtype(dt).itemsize
    assert o == len(raw)
    return out


@pytest.fixture(scope="module")
def grid():
    g = gg.make("grid1k")
    G = bcx.Graph.from_arcs(g.n, g.u, g.v, g.w)
    ix = bcx.Index.build_mhl(G)
    return g, G, ix


def same_answers(a, b, n, pairs=3000, seed=3):
    s = gg.randint(seed, 1, pairs, 0, n - 1)
    t = gg.randint(seed, 2, pairs, 0, n - 1)
    return all(a.query(int(x), int(y)) == b.query(int(x), int(y)) for x, y in zip(s, t))


def test_image_roundtrip_with_h2h(tmp_path, grid):
    g, G, ix = grid
    ix.build_h2h()
    p = tmp_path / "grid.idx"
    ix.save(p)
    back = bcx.Index.load(p)
    assert back.stats()["tree_height"] == ix.stats()["tree_height"]
    assert back.stats()["bag_entries"] == ix.stats()["bag_entries"]
    assert same_answers(ix, back, g.n)
    for s, t in [(0, 1023), (5, 600), (77, 77)]:
        assert back.h2h_query(s, t) == ix.h2h_query(s, t)


def test_truncated_images_are_refused(tmp_path, grid):
    g, G, ix = grid
    p = tmp_path / "grid.idx"
    ix.save(p)
    raw = p.read_bytes()
    for cut in (7, HEAD - 1, HEAD + 5, len(raw) // 3, len(raw) - 1):
        q = tmp_path / f"cut{cut}.idx"
        q.write_bytes(raw[:cut])
        with pytest.raises(bcx.BcxError):
            bcx.Index.load(q)
    q = tmp_path / "long.idx"
    q.write_bytes(raw + b"\0")
    with pytest.raises(bcx.BcxError, match="trailing"):
        bcx.Index.load(q)


def _mutated(tmp_path, raw, fn, name):
    b = bytearray(raw)
    fn(b, vector_offsets(raw))
    q = tmp_path / f"{name}.idx"
    q.write_bytes(bytes(b))
    return q


def _put(b, off, dt, i, value):
    struct.pack_into("<" + {"i4": "i", "i8": "q", "u8": "Q", "u4": "I"}[dt.str[1:]], b, off + i * dt.itemsize, value)


@pytest.mark.parametrize("name,fn,match", [
    ("height", lambda b, o: struct.pack_into("<i", b, 12, 1), "height"),
    ("depth", lambda b, o: _put(b, o["depth"][0], o["depth"][2], 5, 40), "depth"),
    ("parent", lambda b, o: _put(b, o["parent"][0], o["parent"][2], 9, 9), "parent"),
    ("bag-not-ancestor", lambda b, o: _put(b, o["bag"][0], o["bag"][2], o["bag"][1] - 1, 1023), "ancestor"),
    ("bag-offset", lambda b, o: _put(b, o["bag_off"][0], o["bag_off"][2], 3, 10 ** 9), "bag"),
    ("own-entry", lambda b, o: _put(b, o["cnt"][0], o["cnt"][2], 0, 7), "own entry"),
    ("shortcut-length", lambda b, o: _put(b, o["bag_len"][0], o["bag_len"][2], 0, 0), "shortcut length"),
    ("width", lambda b, o: struct.pack_into("<i", b, 16, 3), "width"),
    ("perm", lambda b, o: _put(b, o["perm"][0], o["perm"][2], 1, 0), "permutation"),
])
def test_mutated_images_are_refused(tmp_path, grid, name, fn, match):
    g, G, ix = grid
    p = tmp_path / "grid.idx"
    ix.save(p)
    q = _mutated(tmp_path, p.read_bytes(), fn, name)
    with pytest.raises(bcx.BcxError, match=match):
        bcx.Index.load(q)


def test_p2h_tables_in_the_image(tmp_path, grid):
    g, G, ix = grid
    ix.build_h2h()
    p = tmp_path / "grid.idx"
    ix.save(p)
    raw = p.read_bytes()
    off = vector_offsets(raw)
    assert off["p2h_node"][1] > 0 and off["p2h"][1] > 0
    sep, p2h = ix.h2h_scan_stats(20000, 3)
    assert p2h < sep
    q = _mutated(tmp_path, raw, lambda b, o: _put(b, o["p2h"][0], o["p2h"][2], 0, 10 ** 6), "p2h")
    with pytest.raises(bcx.BcxError, match="P2H"):
        bcx.Index.load(q)


def test_image_of_another_graph_is_refused(tmp_path, grid):
    g, G, ix = grid
    p = tmp_path / "grid.idx"
    ix.save(p)
    other = gg.grid_graph(k=32, seed=2)  # the same 1,024 vertices, other edges
    G2 = bcx.Graph.from_arcs(other.n, other.u, other.v, other.w)
    loaded = bcx.Index.load(p)
    with pytest.raises(bcx.BcxError, match="another graph"):
        bcx.Context(G2, loaded, 0)
    w = g.w.copy()
    w[0] += 1  # one weight changed (both directions, so still a valid graph)
    rev = np.flatnonzero((g.u == g.v[0]) & (g.v == g.u[0]))
    w[rev] += 1
    G3 = bcx.Graph.from_arcs(g.n, g.u, g.v, w)
    with pytest.raises(bcx.BcxError, match="another graph"):
        bcx.Context(G3, loaded, 0)


def test_export_and_from_arrays_roundtrip(grid):
    g, G, ix = grid
    a = ix.export()
    par, hgt = ix.tree()
    assert np.array_equal(a["parent"], par)
    assert np.array_equal(np.diff(a["label_off"]), hgt - 1)
    back = bcx.Index.from_arrays(a, G)
    st, st2 = ix.stats(), back.stats()
    for k in ("tree_height", "tree_width", "label_entries", "bag_entries"):
        assert st[k] == st2[k], k
    assert same_answers(ix, back, g.n)


def test_reference_treedecomp_through_the_seam():
    """The reference's own host index for the grid (tc_ref tcfull: its label
    rows, counts, hop counts and bags) handed over through the host-array
    seam: the tree and the pair answers equal the reference's golden ones."""
    full = dict(np.load(GOLDEN / "grid1k_tcfull.npz"))
    tc = np.load(GOLDEN / "grid1k_tc.npz")
    n = int(tc["n"])
    G = bcx.Graph.from_arcs(n, tc["u"], tc["v"], tc["w"])
    ix = bcx.Index.from_arrays(full, G)
    par, hgt = ix.tree()
    assert np.array_equal(hgt, tc["height"])
    assert np.array_equal(np.where(par < 0, 0, par + 1), tc["parent"])
    for s, t, d, c, h in tc["queries"].tolist():
        assert ix.query(s - 1, t - 1)[:3] == (d, c, h), (s, t)
    own = bcx.Index.build_mhl(G)  # same tree as the reference's, same labels
    assert same_answers(ix, own, n)


def test_from_arrays_refuses_bad_trees(grid):
    g, G, ix = grid
    a = ix.export()
    two_roots = dict(a, parent=np.where(np.arange(g.n) == 7, -1, a["parent"]).astype(np.int32))
    with pytest.raises(bcx.BcxError):
        bcx.Index.from_arrays(two_roots)
    root = int(np.flatnonzero(a["parent"] < 0)[0])
    child = int(np.flatnonzero(a["parent"] == root)[0])
    cyc = a["parent"].copy()
    cyc[root] = child
    with pytest.raises(bcx.BcxError):
        bcx.Index.from_arrays(dict(a, parent=cyc))
    short = dict(a, label_off=a["label_off"].copy())
    short["label_off"][5] += 1
    with pytest.raises(bcx.BcxError):
        bcx.Index.from_arrays(short)
    with pytest.raises(bcx.BcxError, match="different vertex counts"):
        small = gg.make("grid1k", k=10)
        bcx.Index.from_arrays(a, bcx.Graph.from_arcs(small.n, small.u, small.v, small.w))


def test_distance_bound_replaces_weight_sum_cap():
    # A star whose weights sum past the label INF but whose distances stay far
    # below it: accepted (the reference accepts it too).
    n = 5001
    edges = np.array([(0, i) for i in range(1, n)], np.int32)
    star = gg._both_ways(n, edges[:, 0], edges[:, 1], np.full(n - 1, 300_000, np.int64))
    G = bcx.Graph.from_arcs(star.n, star.u, star.v, star.w)
    ix = bcx.Index.build_mhl(G)
    assert ix.query(1, 2)[:2] == (600_000, 1)
    # An arc longer than every path it could shortcut is harmless.
    tri = gg._both_ways(3, np.array([0, 1, 0]), np.array([1, 2, 2]), np.array([5, 5, 2_000_000_000]))
    ix = bcx.Index.build_mhl(bcx.Graph.from_arcs(3, tri.u, tri.v, tri.w))
    assert ix.query(0, 2)[:2] == (10, 1)
    # Distances that can reach the INF are refused at the index build.
    far = gg._both_ways(2, np.array([0]), np.array([1]), np.array([999_999_999]))
    G = bcx.Graph.from_arcs(2, far.u, far.v, far.w)
    with pytest.raises(bcx.BcxError, match="INF"):
        bcx.Index.build_mhl(G)
