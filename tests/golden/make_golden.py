"""Writes the parity fixtures under tests/golden/ (run in the build container).

Inputs: the reference's shipped LosAngeles data (graph, order file, query
batch) and the seeded 32x32 grid workload.  Expected outputs come from
oracle/_ref/tc_ref, the reference's own host index code (TreeDecomp.cpp,
buildIndex.cpp) compiled by oracle/build_ref.sh with a dumping harness:
tree shape (parent, height, label length per vertex), the elimination order
and (distance, count, max-hop) for seeded vertex pairs evaluated with the
reference's scan on the reference's labels.  Fixtures are data only.

  python tests/golden/make_golden.py [--reference /root/reference]
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


def seeded_pairs(n: int, k: int, seed: int) -> np.ndarray:
    a = gg.randint(seed, 1, k, 1, n)
    b = gg.randint(seed, 2, k, 1, n)
    p = np.stack([a, b], axis=1)
    p[:8, 1] = p[:8, 0]  # a few s == t pairs
    return p


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--reference", default="/root/reference")
    args = ap.parse_args()
    ref = Path(args.reference)

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
