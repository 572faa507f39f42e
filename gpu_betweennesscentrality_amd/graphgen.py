"""Seeded synthetic inputs for the betweenness workloads.

All randomness comes from a counter-based splitmix64 stream with explicit
integer mappings, so a (seed, config) pair gives the same graph on every
machine and numpy version.  Graphs are returned as 0-based arc arrays with
every undirected edge listed in both directions (the reference's graph-file
contract, README.md:17-19); ``write_graph`` emits the 1-based text format.

Workloads (vertices, undirected edges m, arc lines A = 2m):
  grid      32x32 grid, 10% of edges deleted keeping connectivity, one random
            diagonal in 5% of the cells, w ~ U[1, 1000]          (seed 1)
  road      n uniform points in the unit square, Delaunay triangulation,
            Euclidean MST plus further Delaunay edges in seeded order until m
            edges, w = round(len * 1e5) + U[0, 10] >= 1; optionally a share of
            edges shortened x0.3 before rounding (arterials).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

_GAMMA = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix(seed: int, stream: int, count: int) -> np.ndarray:
    """count uint64 values of stream `stream` of `seed` (splitmix64 finaliser)."""
    with np.errstate(over="ignore"):
        base = np.uint64((seed * 0x632BE59BD9B4E019 + stream * 0xD1B54A32D192ED03) & (2**64 - 1))
        z = base + (np.arange(1, count + 1, dtype=np.uint64) * _GAMMA)
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        return z ^ (z >> np.uint64(31))


def uniform01(seed: int, stream: int, count: int) -> np.ndarray:
    return (splitmix(seed, stream, count) >> np.uint64(11)).astype(np.float64) * (1.0 / (1 << 53))


def randint(seed: int, stream: int, count: int, lo: int, hi: int) -> np.ndarray:
    """Integers in [lo, hi] (inclusive)."""
    span = np.uint64(hi - lo + 1)
    return (splitmix(seed, stream, count) % span).astype(np.int64) + lo


def permutation(seed: int, stream: int, count: int) -> np.ndarray:
    """Seeded permutation of range(count) (sort by random keys, ties by index)."""
    keys = splitmix(seed, stream, count)
    return np.lexsort((np.arange(count), keys))


@dataclass
class ArcGraph:
    n: int
    u: np.ndarray  # int32, 0-based, both directions
    v: np.ndarray
    w: np.ndarray

    @property
    def arcs(self) -> int:
        return len(self.u)

    @property
    def m(self) -> int:
        return len(self.u) // 2


def _both_ways(n: int, a: np.ndarray, b: np.ndarray, w: np.ndarray) -> ArcGraph:
    """Edges (a, b, w) -> arcs sorted by (tail, head), each edge twice."""
    u = np.concatenate([a, b]).astype(np.int32)
    v = np.concatenate([b, a]).astype(np.int32)
    ww = np.concatenate([w, w]).astype(np.int32)
    o = np.lexsort((v, u))
    return ArcGraph(n, u[o], v[o], ww[o])


class _DSU:
    def __init__(self, n):
        self.p = np.arange(n)

    def find(self, x):
        p = self.p
        r = x
        while p[r] != r:
            r = p[r]
        while p[x] != r:
            p[x], x = r, p[x]
        return r

    def union(self, a, b):
        ra, rb = self.find(a), self.find(b)
        if ra == rb:
            return False
        self.p[ra] = rb
        return True


def _connected(n: int, a: np.ndarray, b: np.ndarray) -> bool:
    d = _DSU(n)
    comps = n
    for x, y in zip(a.tolist(), b.tolist()):
        if d.union(x, y):
            comps -= 1
    return comps == 1


def grid_graph(k: int = 32, seed: int = 1, delete_frac: float = 0.10,
               diag_frac: float = 0.05, wmax: int = 1000, unit: bool = False) -> ArcGraph:
    """k x k grid with seeded deletions (connectivity kept) and diagonals."""
    n = k * k
    edges = []
    for r in range(k):
        for c in range(k):
            if c + 1 < k:
                edges.append((r * k + c, r * k + c + 1))
            if r + 1 < k:
                edges.append((r * k + c, (r + 1) * k + c))
    E = np.array(edges, dtype=np.int64)
    keep = np.ones(len(E), bool)
    target = int(round(delete_frac * len(E)))
    deleted = 0
    for e in permutation(seed, 1, len(E)).tolist():
        if deleted >= target:
            break
        keep[e] = False
        if _connected(n, E[keep, 0], E[keep, 1]):
            deleted += 1
        else:
            keep[e] = True
    E = E[keep]
    cells = (k - 1) * (k - 1)
    nd = int(round(diag_frac * cells))
    chosen = permutation(seed, 2, cells)[:nd]
    orient = randint(seed, 3, nd, 0, 1)
    diags = []
    for cell, o in zip(chosen.tolist(), orient.tolist()):
        r, c = divmod(cell, k - 1)
        if o == 0:
            diags.append((r * k + c, (r + 1) * k + c + 1))
        else:
            diags.append((r * k + c + 1, (r + 1) * k + c))
    if diags:
        E = np.concatenate([E, np.array(diags, dtype=np.int64)])
    w = np.ones(len(E), np.int64) if unit else randint(seed, 4, len(E), 1, wmax)
    return _both_ways(n, E[:, 0], E[:, 1], w)


def unit_grid(k: int) -> ArcGraph:
    """Plain k x k unit grid (closed-form path counts)."""
    n = k * k
    a, b = [], []
    for r in range(k):
        for c in range(k):
            if c + 1 < k:
                a.append(r * k + c)
                b.append(r * k + c + 1)
            if r + 1 < k:
                a.append(r * k + c)
                b.append((r + 1) * k + c)
    a = np.array(a)
    return _both_ways(n, a, np.array(b), np.ones(len(a), np.int64))


def road_graph(n: int, m: int, seed: int, arterial_frac: float = 0.0) -> ArcGraph:
    """Road-like planar graph: Delaunay -> MST + seeded extra edges up to m."""
    from scipy.sparse import coo_matrix
    from scipy.sparse.csgraph import minimum_spanning_tree
    from scipy.spatial import Delaunay

    xy = np.stack([uniform01(seed, 10, n), uniform01(seed, 11, n)], axis=1)
    tri = Delaunay(xy).simplices.astype(np.int64)
    e = np.concatenate([tri[:, [0, 1]], tri[:, [1, 2]], tri[:, [0, 2]]])
    e.sort(axis=1)
    e = np.unique(e, axis=0)
    if m > len(e):
        raise ValueError(f"Delaunay has only {len(e)} edges; asked for {m}")
    length = np.hypot(*(xy[e[:, 0]] - xy[e[:, 1]]).T)
    mst = minimum_spanning_tree(coo_matrix((length, (e[:, 0], e[:, 1])), shape=(n, n))).tocoo()
    t = np.stack([mst.row, mst.col], axis=1).astype(np.int64)
    t.sort(axis=1)
    key = e[:, 0] * n + e[:, 1]
    in_tree = np.isin(key, t[:, 0] * n + t[:, 1])
    if in_tree.sum() != n - 1:
        raise RuntimeError("MST is not spanning")
    rest = np.flatnonzero(~in_tree)
    rest = rest[permutation(seed, 12, len(rest))][: m - (n - 1)]
    sel = np.sort(np.concatenate([np.flatnonzero(in_tree), rest]))
    e, length = e[sel], length[sel]
    if arterial_frac > 0:
        k = int(round(arterial_frac * len(e)))
        idx = permutation(seed, 13, len(e))[:k]
        length = length.copy()
        length[idx] *= 0.3
    w = np.rint(length * 1e5).astype(np.int64) + randint(seed, 14, len(e), 0, 10)
    w = np.maximum(w, 1)
    return _both_ways(n, e[:, 0], e[:, 1], w)


# Named workloads.  Edge counts follow the DIMACS road-network convention of
# the sizes they mirror (BAY 321,270 / 800,172 arcs, FLA 1,070,376 / 2,712,798
# arcs): about 2.5 arc lines per vertex, i.e. mean degree 2.5.
WORKLOADS = {
    "grid1k": dict(kind="grid", k=32, seed=1),
    "road100k": dict(kind="road", n=100_000, m=125_000, seed=2),
    "road320k": dict(kind="road", n=320_000, m=400_000, seed=3),
    "road1070k": dict(kind="road", n=1_070_000, m=1_350_000, seed=5, arterial_frac=0.01),
    # The other reading of the configs' edge counts: undirected edges, 2.5 per
    # vertex (1.6M arc lines at 320k) -- a near-triangulation, far denser
    # than road networks (tree height 2,368 against 526).
    "road320k_m800k": dict(kind="road", n=320_000, m=800_000, seed=3),
    # The same reading at 100k vertices: 250,000 undirected edges (tree
    # height 1,235 against 294).
    "road100k_m250k": dict(kind="road", n=100_000, m=250_000, seed=2),
}


def make(name: str, **override) -> ArcGraph:
    spec = dict(WORKLOADS[name])
    spec.update(override)
    kind = spec.pop("kind")
    if kind == "grid":
        return grid_graph(**spec)
    return road_graph(**spec)


def write_graph(g: ArcGraph, path) -> None:
    """Reference text format: "n A" then A lines "u v w" (1-based)."""
    with open(path, "w") as f:
        f.write(f"{g.n} {g.arcs}\n")
        np.savetxt(f, np.stack([g.u + 1, g.v + 1, g.w], axis=1), fmt="%d")


def write_queries(ids, path) -> None:
    """RQ query file: "Q" then Q 1-based ids (RQ.cu:589-610)."""
    ids = np.asarray(ids)
    with open(path, "w") as f:
        f.write(f"{len(ids)}\n")
        np.savetxt(f, ids + 1, fmt="%d")


def query_sets(n: int, count: int, seed: int, qmin: int = 1000, qmax: int = 50_000,
               adjacency=None) -> list[np.ndarray]:
    """RQ request sets: Q log-uniform in [qmin, qmax]; uniform-random distinct
    vertices or a hop-BFS ball around a random centre, with equal probability."""
    u = uniform01(seed, 20, count)
    qs = np.rint(np.exp(np.log(qmin) + u * (np.log(qmax) - np.log(qmin)))).astype(np.int64)
    kinds = randint(seed, 21, count, 0, 1)
    out = []
    for i, (q, kind) in enumerate(zip(qs.tolist(), kinds.tolist())):
        q = min(q, n)
        if kind == 0 or adjacency is None:
            out.append(np.sort(permutation(seed, 100 + i, n)[:q]).astype(np.int32))
        else:
            row, col = adjacency
            centre = int(randint(seed, 200 + i, 1, 0, n - 1)[0])
            seen = {centre}
            frontier = [centre]
            order = [centre]
            while frontier and len(order) < q:
                nxt = []
                for x in frontier:
                    for y in col[row[x]:row[x + 1]].tolist():
                        if y not in seen:
                            seen.add(y)
                            nxt.append(y)
                            order.append(y)
                frontier = nxt
            out.append(np.sort(np.array(order[:q], dtype=np.int32)))
    return out
