"""Writes the parity fixtures under tests/golden/ (run in the build container).

Inputs: the reference's shipped LosAngeles data (graph, order file, query
batch) and the seeded 32x32 grid workload.  Expected outputs come from
oracle/_ref/tc_ref, the reference's own host index code (TreeDecomp.cpp,
buildIndex.cpp) compiled by oracle/build_ref.sh with a dumping harness:
tree shape (parent, height, label length per vertex), the elimination order
and (distance, count, max-hop) for seeded vertex pairs evaluated with the
reference's scan on the reference's labels; for the grid also the reference's
whole host index structures, the inputs of the host-array seams: TreeDecomp's
label rows and bags, and QBC's H2H distance rows, optimizeVS separators and P2H
branch tables.  Fixtures are data only.

  python tests/golden/make_golden.py [--reference /root/reference] [--only-h2hfull]
"""
from __future__ import annotations

import argparse
import subprocess
import sys
import tempfile
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent.parent
sys.path.insert(0, str(ROOT))

from gpu_betweennesscentrality_amd import graphgen as gg  # noqa: E402


def parse_dump(text: str, n: int):
    parent = np.zeros(n, np.int32)
    height = np.zeros(n, np.int32)
    rowlen = np.zeros(n, np.int32)
    order = None
    q = []
    stats = None
    for line in text.splitlines():
        f = line.split()
        if not f:
            continue
        if f[0] == "n" and len(f) >= 8:
            stats = dict(n=int(f[1]), height=int(f[3]), width=int(f[5]), entries=int(f[7]))
        elif f[0] == "order":
            order = np.array([int(x) for x in f[1:]], np.int32)
        elif f[0] == "V":
            v = int(f[1]) - 1
            parent[v] = int(f[2])
            height[v] = int(f[3])
            rowlen[v] = int(f[4])
        elif f[0] == "Q":
            q.append([int(f[1]), int(f[2]), int(f[3]), int(f[4]), int(f[5])])
    return stats, parent, height, rowlen, order, np.array(q, np.int64).reshape(-1, 5)


def run_ref(args, pairs: np.ndarray, n: int):
    exe = ROOT / "oracle" / "_ref" / "tc_ref"
    if not exe.exists():
        raise SystemExit("oracle/_ref/tc_ref missing: run sh oracle/build_ref.sh")
    with tempfile.TemporaryDirectory() as d:
        pp = Path(d) / "pairs.txt"
        np.savetxt(pp, pairs, fmt="%d")
        out = subprocess.run([str(exe), *map(str, args), str(pp)], capture_output=True, text=True,
                             check=True).stdout
    return parse_dump(out, n)


def run_ref_full(graph_path, n: int) -> dict:
    """The reference TreeDecomp's whole host structure (tc_ref tcfull):
    parent, label rows over ancestors and bags, indexed by 0-based vertex id."""
    exe = ROOT / "oracle" / "_ref" / "tc_ref"
    out = subprocess.run([str(exe), "tcfull", str(graph_path)], capture_output=True, text=True,
                         check=True).stdout
    parent = np.full(n, -1, np.int32)
    rows = [None] * n
    bags = [None] * n
    for line in out.splitlines():
        f = line.split()
        if not f:
            continue
        if f[0] == "V":
            parent[int(f[1]) - 1] = int(f[2]) - 1
        elif f[0] == "R":
            rows[int(f[1]) - 1] = [tuple(int(x) for x in e.split(":")) for e in f[3:]]
        elif f[0] == "B":
            bags[int(f[1]) - 1] = [tuple(int(x) for x in e.split(":")) for e in f[3:]]
    lo = np.zeros(n + 1, np.int64)
    bo = np.zeros(n + 1, np.int64)
    for v in range(n):
        lo[v + 1] = lo[v] + len(rows[v])
        bo[v + 1] = bo[v] + len(bags[v])
    R = np.array([e for r in rows for e in r], np.int64).reshape(-1, 3)
    Bg = np.array([e for b in bags for e in b], np.int64).reshape(-1, 4)
    return dict(parent=parent, label_off=lo, dis=R[:, 0].astype(np.int32), cnt=R[:, 1].astype(np.uint64),
                hop=R[:, 2].astype(np.uint32), bag_off=bo, bag_v=(Bg[:, 0] - 1).astype(np.int32),
                bag_len=Bg[:, 1].astype(np.int32), bag_cnt=Bg[:, 2].astype(np.uint64),
                bag_hop=Bg[:, 3].astype(np.uint32))


def run_ref_h2hfull(graph_path, order_path, n: int) -> dict:
    """QBC's whole host H2H index (tc_ref h2hfull): parent, distance rows over
    ancestors, optimizeVS separators and the kept P2H branch tables, indexed
    by 0-based vertex id (the layout of bcx_index_h2h_from_arrays)."""
    exe = ROOT / "oracle" / "_ref" / "tc_ref"
    out = subprocess.run([str(exe), "h2hfull", str(graph_path), str(order_path)], capture_output=True,
                         text=True, check=True).stdout
    parent = np.full(n, -1, np.int32)
    rows = [None] * n
    seps = [None] * n
    lists = {}
    for line in out.splitlines():
        f = line.split()
        if not f:
            continue
        if f[0] == "V":
            parent[int(f[1]) - 1] = int(f[2]) - 1
        elif f[0] == "D":
            rows[int(f[1]) - 1] = [int(x) for x in f[3:]]
        elif f[0] == "S":
            seps[int(f[1]) - 1] = [int(x) for x in f[3:]]
        elif f[0] == "L":
            lists.setdefault(int(f[1]) - 1, {})[int(f[2]) - 1] = [int(x) for x in f[4:]]
    go = np.zeros(n + 1, np.int64)
    so = np.zeros(n + 1, np.int64)
    for v in range(n):
        go[v + 1] = go[v] + len(rows[v])
        so[v + 1] = so[v] + len(seps[v])
    branch = np.array(sorted(lists), np.int32)
    po = np.zeros(len(branch) * n + 1, np.int64)
    flat = []
    for j, b in enumerate(branch.tolist()):
        for v in range(n):
            po[j * n + v] = len(flat)
            flat += lists[b].get(v, [])
    po[-1] = len(flat)
    return dict(parent=parent, gdis_off=go, gdis=np.array([x for r in rows for x in r], np.int32),
                sep_off=so, sep=np.array([x for r in seps for x in r], np.int32), branch=branch,
                p2h_off=po, p2h=np.array(flat, np.int32))


def grid_h2hfull() -> None:
    """grid1k_h2hfull.npz: the reference's QBC H2H index of the grid on its
    min-degree order (the order GraphConstract writes, restated by
    bcx_min_degree_order and checked against the shipped LosAngeles order)."""
    import gpu_betweennesscentrality_amd as bcx

    g = gg.make("grid1k")
    order = bcx.Graph.from_arcs(g.n, g.u, g.v, g.w).min_degree_order()
    with tempfile.TemporaryDirectory() as d:
        gp, op = Path(d) / "grid.txt", Path(d) / "order.txt"
        gg.write_graph(g, gp)
        with open(op, "w") as f:
            f.write(f"{g.n}\n")
            for v in range(g.n):
                f.write(f"{v} {int(order[v])}\n")
        full = run_ref_h2hfull(gp, op, g.n)
    np.savez_compressed(HERE / "grid1k_h2hfull.npz", order=order.astype(np.int32), **full)
    print("grid1k full H2H:", len(full["gdis"]), "distance entries,", len(full["sep"]), "separator entries,",
          len(full["branch"]), "branch vertices,", len(full["p2h"]), "P2H entries")


def seeded_pairs(n: int, k: int, seed: int) -> np.ndarray:
    a = gg.randint(seed, 1, k, 1, n)
    b = gg.randint(seed, 2, k, 1, n)
    p = np.stack([a, b], axis=1)
    p[:8, 1] = p[:8, 0]  # a few s == t pairs
    return p


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--reference", default="/root/reference")
    ap.add_argument("--only-h2hfull", action="store_true", help="write grid1k_h2hfull.npz only")
    args = ap.parse_args()
    ref = Path(args.reference)
    grid_h2hfull()
    if args.only_h2hfull:
        return

    # Seeded grid workload.
    g = gg.make("grid1k")
    with tempfile.TemporaryDirectory() as d:
        gp = Path(d) / "grid.txt"
        gg.write_graph(g, gp)
        pairs = seeded_pairs(g.n, 3000, 101)
        st, parent, height, rowlen, order, q = run_ref(["tc", gp], pairs, g.n)
    np.savez_compressed(HERE / "grid1k_tc.npz", u=g.u, v=g.v, w=g.w, n=g.n, parent=parent,
                        height=height, rowlen=rowlen, order=order, queries=q,
                        stats=np.array([st["height"], st["width"], st["entries"]]))
    print("grid1k:", st, len(q), "queries")
    with tempfile.TemporaryDirectory() as d:
        gp = Path(d) / "grid.txt"
        gg.write_graph(g, gp)
        full = run_ref_full(gp, g.n)
    np.savez_compressed(HERE / "grid1k_tcfull.npz", **full)
    print("grid1k full TreeDecomp:", len(full["dis"]), "label entries,", len(full["bag_v"]), "bag entries")

    # LosAngeles fixture shipped with the reference.
    gpath = ref / "data" / "roadData" / "LosAngeles.txt"
    opath = ref / "data" / "orderFile" / "LosAngelesOrder.txt"
    qpath = ref / "data" / "queryBatch" / "LosAngeles.txt"
    raw = np.loadtxt(gpath, dtype=np.int64, skiprows=1)
    n = int(open(gpath).readline().split()[0])
    ordr = np.loadtxt(opath, dtype=np.int64, skiprows=1)
    order = np.zeros(n, np.int32)
    order[ordr[:, 0]] = ordr[:, 1]
    qids = np.loadtxt(qpath, dtype=np.int64, skiprows=1).astype(np.int32)
    np.savez_compressed(HERE / "la_graph.npz", n=n, u=(raw[:, 0] - 1).astype(np.int32),
                        v=(raw[:, 1] - 1).astype(np.int32), w=raw[:, 2].astype(np.int32),
                        order=order, queries=qids - 1)
    pairs = seeded_pairs(n, 2000, 202)
    st, parent, height, rowlen, elim, q = run_ref(["tc", gpath], pairs, n)
    np.savez_compressed(HERE / "la_tc.npz", parent=parent, height=height, rowlen=rowlen,
                        order=elim, queries=q,
                        stats=np.array([st["height"], st["width"], st["entries"]]))
    print("LA tc:", st, len(q), "queries")
    st, parent, height, rowlen, _, q = run_ref(["h2h", gpath, opath], pairs, n)
    np.savez_compressed(HERE / "la_h2h.npz", parent=parent, height=height, rowlen=rowlen,
                        queries=q[:, :3], stats=np.array([st["height"], st["entries"]]))
    print("LA h2h:", st, len(q), "queries")


if __name__ == "__main__":
    main()
