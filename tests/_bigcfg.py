"""Shared helpers of the parity tests at the benchmark configurations' sizes."""
from concurrent.futures import ThreadPoolExecutor

import numpy as np

import gpu_betweennesscentrality_amd as bcx
from gpu_betweennesscentrality_amd import graphgen as gg
import oracle

BC_RTOL = 1e-9


def close_bc(a, b, rtol=BC_RTOL):
    a, b = np.asarray(a), np.asarray(b)
    return bool(np.all(np.abs(a - b) <= rtol * np.maximum(np.abs(b), 1.0)))


def max_rel(a, b):
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), 1.0)))


def workload(name):
    print(f"[cfg] building {name}", flush=True)
    g = gg.make(name)
    G = bcx.Graph.from_arcs(g.n, g.u, g.v, g.w)
    ix = bcx.Index.build_mhl(G)
    ctx = bcx.Context(G, ix, 0)
    print(f"[cfg] {name}: n={g.n} arcs={g.arcs} {ix.stats()}", flush=True)
    return g, oracle.CSR.from_arcgraph(g), G, ix, ctx


def sample_sources(ix, n, consecutive, scattered, seed):
    """`consecutive` sources from the middle of the tree order (one or more
    whole tiles of neighbours) and `scattered` seeded random ones."""
    tree = ix.tree_order()
    mid = (n // 2) - (n // 2) % 64
    a = tree[mid:mid + consecutive]
    b = gg.permutation(seed, 1, n)[:scattered]
    return np.concatenate([a, b]).astype(np.int32)


def oracle_rows(csr, sources):
    """Oracle (d, sigma, hop) rows, one Dijkstra per source on a thread pool
    (the oracle releases the interpreter lock)."""
    with ThreadPoolExecutor(oracle.default_threads()) as ex:
        return list(ex.map(lambda s: oracle.sssp(csr, int(s)), list(sources)))


def rows_equal(got, want_rows, hop=True):
    d, s, h, _ = got
    for i, (od, osg, ohp, _) in enumerate(want_rows):
        if not np.array_equal(d[i].astype(np.int64), od):
            return f"distances of row {i}"
        if not np.array_equal(s[i], osg):
            return f"counts of row {i}"
        if hop and h is not None and not np.array_equal(h[i], ohp):
            return f"layers of row {i}"
    return None


def widest_level(ix):
    _, height = ix.tree()
    return int(np.bincount(height).max())
