"""GPU parity: the HIP path (through the C ABI) against the CPU oracle.

Bars: distances, uint64 path counts and max-hop layers bit-exact; BC within
1e-9 relative (|a - b| <= 1e-9 * max(|b|, 1)); plain and batched forms
bitwise identical (same accumulation order by construction).
"""
import json
import os
import subprocess
from pathlib import Path

import numpy as np
import pytest

import gpu_betweennesscentrality_amd as bcx
from gpu_betweennesscentrality_amd import graphgen as gg
import oracle

pytestmark = pytest.mark.gpu
ROOT = Path(__file__).resolve().parent.parent
BC_RTOL = 1e-9


def close_bc(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return np.all(np.abs(a - b) <= BC_RTOL * np.maximum(np.abs(b), 1.0))


def build(g):
    G = bcx.Graph.from_arcs(g.n, g.u, g.v, g.w)
    ix = bcx.Index.build_mhl(G)
    return G, ix, bcx.Context(G, ix, 0)


@pytest.fixture(scope="module")
def grid():
    g = gg.make("grid1k")
    G, ix, ctx = build(g)
    return g, oracle.CSR.from_arcgraph(g), ctx


def test_query_grid_all_sources(grid):
    g, csr, ctx = grid
    for s in range(g.n):
        d, sg, hp, ovf = ctx.query(s)
        od, osg, ohp, oovf = oracle.sssp(csr, s)
        assert not ovf and not oovf
        assert np.array_equal(d.astype(np.int64), od), f"distance mismatch from source {s}"
        assert np.array_equal(sg, osg), f"count mismatch from source {s}"
        assert np.array_equal(hp, ohp), f"hop mismatch from source {s}"


def test_full_bc_grid_plain_and_batched(grid):
    g, csr, ctx = grid
    ref, _ = oracle.brandes(csr)
    plain, o1 = ctx.full_bc(bcx.MODE_PLAIN)
    ctx.set_tuning(batch=256)
    batched, o2 = ctx.full_bc(bcx.MODE_BATCHED)
    assert not o1 and not o2
    assert close_bc(plain, ref)
    assert close_bc(batched, ref)
    assert np.array_equal(plain.view(np.uint64), batched.view(np.uint64))


@pytest.mark.parametrize("scale", [1, 4000])
def test_label_scan_small_and_large_distances(scale):
    """The label-scan tile kernel on the seeded grid and on the same grid
    with weights times 4,000 (distances near 2^28, sums of two label entries
    near 2^29): rows bit-exact against the oracle from every source, and
    full BC bitwise equal to the one-source-per-launch form."""
    g = gg.make("grid1k")
    g = gg.ArcGraph(g.n, g.u, g.v, (g.w.astype(np.int64) * scale).astype(np.int32))
    csr = oracle.CSR.from_arcgraph(g)
    G, ix, ctx = build(g)
    src = np.arange(g.n, dtype=np.int32)
    ctx.set_tuning(batch=256)
    d, s, h, _ = ctx.query_rows(src, bcx.MODE_BATCHED)
    for i in range(g.n):
        od, osg, ohp, _ = oracle.sssp(csr, i)
        assert np.array_equal(d[i].astype(np.int64), od), f"scale {scale}, distances of row {i}"
        assert np.array_equal(s[i], osg) and np.array_equal(h[i], ohp), f"scale {scale}, row {i}"
    batched, o1 = ctx.full_bc(bcx.MODE_BATCHED)
    plain, o2 = ctx.full_bc(bcx.MODE_PLAIN)
    assert not o1 and not o2
    assert np.array_equal(batched.view(np.uint64), plain.view(np.uint64))
    ref, _ = oracle.brandes(csr)
    assert close_bc(batched, ref)
    ctx.close()


@pytest.mark.parametrize("batch", [1, 64, 100, 4096])
def test_batch_size_invariance(grid, batch):
    g, csr, ctx = grid
    src = gg.permutation(7, 1, g.n)[:300].astype(np.int32)
    ctx.set_tuning(batch=batch)
    out, _ = ctx.bc_sources(src, bcx.MODE_BATCHED)
    ref, _ = oracle.brandes(csr, src)
    assert close_bc(out, ref)
    ctx.set_tuning(batch=4096)
    base, _ = ctx.bc_sources(src, bcx.MODE_PLAIN)
    assert np.array_equal(out.view(np.uint64), base.view(np.uint64))


def test_unit_grid_counts_and_overflow():
    from math import comb

    g = gg.unit_grid(34)
    G, ix, ctx = build(g)
    d, sg, hp, ovf = ctx.query(0)
    k = 34
    assert not ovf
    assert int(sg[k * k - 1]) == comb(2 * k - 2, k - 1)
    assert int(d[k * k - 1]) == 2 * k - 2 and int(hp[k * k - 1]) == 2 * k - 2
    g = gg.unit_grid(35)
    G, ix, ctx = build(g)
    d, sg, hp, ovf = ctx.query(0)
    assert ovf, "C(68, 34) > 2^64 must raise the overflow flag"


@pytest.fixture(scope="module")
def road():
    g = gg.road_graph(20_000, 25_000, seed=11)
    G, ix, ctx = build(g)
    return g, oracle.CSR.from_arcgraph(g), ctx


def test_road_queries_sampled(road):
    g, csr, ctx = road
    for s in gg.permutation(3, 1, g.n)[:24].tolist():
        d, sg, hp, _ = ctx.query(s)
        od, osg, ohp, _ = oracle.sssp(csr, s)
        assert np.array_equal(d.astype(np.int64), od)
        assert np.array_equal(sg, osg)
        assert np.array_equal(hp, ohp)


def test_road_bc_sources(road):
    g, csr, ctx = road
    src = gg.permutation(5, 2, g.n)[:512].astype(np.int32)
    ctx.set_tuning(batch=128)
    out, ovf = ctx.bc_sources(src, bcx.MODE_BATCHED)
    ref, _ = oracle.brandes(csr, src)
    assert not ovf
    assert close_bc(out, ref)


def test_subset_bc_matches_full(grid):
    g, csr, ctx = grid
    ref, _ = oracle.brandes(csr)
    pts = np.array([0, 5, 77, 500, 1023], np.int32)
    out = ctx.subset_bc(pts)
    assert close_bc(out, ref[pts])


@pytest.mark.parametrize("mode", [bcx.MODE_SWEEP, bcx.MODE_BATCHED])
def test_rq_engine_alone_and_mixed(grid, mode):
    g, csr, ctx = grid
    ref, _ = oracle.brandes(csr)
    ctx.set_tuning(batch=128, sweep_batch=128)
    a = np.array([3, 9, 400, 1000], np.int32)
    b = np.arange(0, 1024, 7, dtype=np.int32)
    srv = bcx.RQServer(ctx, 1 << 30, mode)
    ra = srv.submit(a)
    alone = None
    while alone is None:
        srv.step()
        alone = srv.poll(ra)
    srv2 = bcx.RQServer(ctx, 1 << 30, mode)
    rb = srv2.submit(b)
    for _ in range(3):
        srv2.step()
    ra2 = srv2.submit(a)  # joins mid-sweep
    res = {}
    while len(res) < 2:
        srv2.step()
        for rid in (ra2, rb):
            if rid not in res:
                r = srv2.poll(rid)
                if r is not None:
                    res[rid] = r
    assert close_bc(alone, ref[a])
    assert close_bc(res[rb], ref[b])
    assert np.array_equal(alone.view(np.uint64), res[ra2].view(np.uint64))


def test_cli_mhlqbc_and_qbc(tmp_path, grid):
    g, csr, ctx = grid
    ref, _ = oracle.brandes(csr)
    gp = tmp_path / "g.txt"
    gg.write_graph(g, gp)
    exe = ROOT / "gpu_betweennesscentrality_amd" / "bin"
    env = dict(os.environ, BCX_BC_DUMP=str(tmp_path / "bc.bin"))
    out = subprocess.run([exe / "MHLQBC", gp, "0", tmp_path / "bc.txt"], env=env,
                         capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    assert "seconds (---totalTime---)" in out.stdout
    dump = np.fromfile(tmp_path / "bc.bin", np.float64)
    assert close_bc(dump, ref)
    lines = (tmp_path / "bc.txt").read_text().splitlines()
    assert lines[0] == "vertexID  BCValue"
    assert lines[1] == f"1 {ref[0]:.2f}"
    # QBC with an order file: contraction order = vertex id order.
    order = tmp_path / "order.txt"
    order.write_text(f"{g.n}\n" + "".join(f"{v} {v}\n" for v in range(g.n)))
    check = tmp_path / "check.txt"
    check.write_text(f"{g.n} {g.arcs}\n" + "".join(f"{v + 1} {ref[v]:.6f}\n" for v in range(g.n)))
    out = subprocess.run([exe / "QBC", gp, order, "0", tmp_path / "bc2.txt", check], env=env,
                         capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "BC value is correct" in out.stdout
    assert close_bc(np.fromfile(tmp_path / "bc.bin", np.float64), ref)


def test_cli_source_ranges_and_metrics(tmp_path, grid):
    """A run split into source ranges (BCX_SOURCES, for resuming a long run
    range by range): each range's dump equals the oracle over those sources,
    and the ranges' dumps sum to the full BC.  BCX_METRICS appends one JSON
    line per run (the reference's stdout is unchanged)."""
    g, csr, ctx = grid
    ref, _ = oracle.brandes(csr)
    gp = tmp_path / "g.txt"
    gg.write_graph(g, gp)
    exe = ROOT / "gpu_betweennesscentrality_amd" / "bin"
    metrics = tmp_path / "metrics.jsonl"
    total = np.zeros(g.n)
    for a, b in ((0, 300), (300, 301), (301, g.n)):
        dump = tmp_path / f"bc_{a}.bin"
        env = dict(os.environ, BCX_SOURCES=f"{a}:{b}", BCX_BC_DUMP=str(dump), BCX_METRICS=str(metrics))
        out = subprocess.run([exe / "MHLQBC", gp, "0"], env=env, capture_output=True, text=True, timeout=120)
        assert out.returncode == 0, out.stderr
        part = np.fromfile(dump, np.float64)
        want, _ = oracle.brandes(csr, np.arange(a, b, dtype=np.int32))
        assert close_bc(part, want), f"range {a}:{b}"
        total += part
    assert close_bc(total, ref)
    qp = tmp_path / "q.txt"
    qp.write_text("3\n1\n2\n3\n")
    env = dict(os.environ, BCX_METRICS=str(metrics))
    out = subprocess.run([exe / "RQ", gp, qp, "0"], env=env, capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    env = dict(os.environ, BCX_METRICS=str(metrics), BCX_PROFILE="1", BCX_BC_DUMP=str(tmp_path / "prof.bin"))
    out = subprocess.run([exe / "MHLQBC", gp, "0"], env=env, capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    assert close_bc(np.fromfile(tmp_path / "prof.bin", np.float64), ref)
    rows = [json.loads(line) for line in metrics.read_text().splitlines()]
    assert rows[4]["tool"] == "MHLQBC" and rows[4]["kernel_ms"]["scan"] > 0  # kernel times under BCX_PROFILE
    rows = rows[:4]
    assert [r["tool"] for r in rows] == ["MHLQBC"] * 3 + ["RQ"]
    assert [r["sources"] for r in rows[:3]] == [300, 1, g.n - 301]
    assert rows[3]["queries"] == 3 and rows[3]["n"] == g.n and rows[3]["overflow"] is False
    for r in rows[:3]:
        assert r["mode"] == "sweep" and r["compute_s"] > 0 and r["sources_per_s"] > 0
        assert abs(r["total_s"] - (r["index_s"] + r["context_s"] + r["compute_s"])) < 1e-3


def test_cli_rq_table(tmp_path, grid):
    """RQ executable: query-file contract and stdout table (RQ.cu:542-610)."""
    g, csr, ctx = grid
    ref, _ = oracle.brandes(csr)
    gp = tmp_path / "g.txt"
    gg.write_graph(g, gp)
    ids = [1, 2, 17, 300, 512, 1024]  # 1-based, as in the reference's query files
    qp = tmp_path / "q.txt"
    qp.write_text(f"{len(ids)}\n" + "\n".join(map(str, ids)) + "\n")
    exe = ROOT / "gpu_betweennesscentrality_amd" / "bin" / "RQ"
    out = subprocess.run([exe, gp, qp, "0"], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    assert f"The number of vertices to query: {len(ids)}" in out.stdout
    lines = out.stdout.splitlines()
    k = lines.index("-" * 20)
    rows = [ln.split() for ln in lines[k + 1:k + 1 + len(ids)]]
    assert [int(r[0]) for r in rows] == ids
    got = np.array([float(r[1]) for r in rows])
    want = ref[np.array(ids) - 1]
    assert np.all(np.abs(got - want) <= 1e-9 * np.maximum(np.abs(want), 1.0) + 5e-9)


# ---- the reference's LosAngeles data ------------------------------------------

def test_la_queries_and_bc_on_the_gpu():
    """LosAngeles (49,374 vertices): the GPU's distance/count/max-hop rows
    against the answers of the reference's own index code (golden pairs),
    and BC of sampled sources in both forms against the oracle."""
    golden = ROOT / "tests" / "golden"
    la = np.load(golden / "la_graph.npz")
    ref = np.load(golden / "la_tc.npz")
    n = int(la["n"])
    G = bcx.Graph.from_arcs(n, la["u"], la["v"], la["w"])
    ix = bcx.Index.build_mhl(G)
    ctx = bcx.Context(G, ix, 0)
    q = ref["queries"].astype(np.int64)
    for s in np.unique(q[:, 0]).tolist():
        d, sg, hp, _ = ctx.query(s - 1)
        rows = q[q[:, 0] == s]
        t = rows[:, 1] - 1
        assert np.array_equal(d[t].astype(np.int64), rows[:, 2]), f"distances from {s}"
        assert np.array_equal(sg[t].astype(np.int64), rows[:, 3]), f"counts from {s}"
        assert np.array_equal(hp[t].astype(np.int64), rows[:, 4]), f"layers from {s}"
    csr = oracle.CSR(n, la["u"].astype(np.int64), la["v"].astype(np.int64), la["w"].astype(np.int64))
    src = gg.permutation(21, 1, n)[:192].astype(np.int32)
    want, _ = oracle.brandes(csr, src)
    for mode in (bcx.MODE_SWEEP, bcx.MODE_BATCHED):
        got, ovf = ctx.bc_sources(src, mode)
        assert not ovf and close_bc(got, want), f"mode {mode}"


# ---- closed forms and edge cases, every form -----------------------------------

def _small(n, edges):
    e = np.asarray(edges, np.int32).reshape(-1, 2)
    g = gg._both_ways(n, e[:, 0], e[:, 1], np.ones(len(e), np.int32))
    G = bcx.Graph.from_arcs(g.n, g.u, g.v, g.w)
    ix = bcx.Index.build_mhl(G)
    ix.build_h2h()
    return g, bcx.Context(G, ix, 0)


FORMS = [bcx.MODE_SWEEP, bcx.MODE_BATCHED, bcx.MODE_PLAIN]


@pytest.mark.parametrize("name,n,edges,want", [
    ("path", 50, [(i, i + 1) for i in range(49)], [i * (49 - i) for i in range(50)]),
    ("star", 40, [(0, i) for i in range(1, 40)], [38 * 39 / 2] + [0] * 39),
    ("odd cycle", 31, [(i, (i + 1) % 31) for i in range(31)], [30 * 28 / 8] * 31),
    ("even cycle", 20, [(i, (i + 1) % 20) for i in range(20)], [18 * 18 / 8] * 20),
    ("complete", 12, [(i, j) for i in range(12) for j in range(i + 1, 12)], [0] * 12),
    ("two vertices", 2, [(0, 1)], [0, 0]),
])
def test_closed_forms_every_form(name, n, edges, want):
    g, ctx = _small(n, edges)
    for mode in FORMS:
        bc, ovf = ctx.full_bc(mode)
        assert not ovf
        assert close_bc(bc, np.asarray(want, np.float64)), f"{name}, mode {mode}"
    csr = oracle.CSR.from_arcgraph(g)
    dist = ctx.h2h_distances(np.arange(n, dtype=np.int32))
    for s in range(n):
        assert np.array_equal(dist[s].astype(np.int64), oracle.sssp(csr, s)[0]), f"{name}, source {s}"


def _random_small_graph(seed):
    """Seeded connected graph of 2..200 vertices: a random spanning tree plus
    up to 2n extra edges, weights from a small range (many equal-length
    paths) or a wide one, and a random contraction order."""
    rng = np.random.default_rng(seed)
    n = int(rng.integers(2, 201))
    edges = {(min(v, int(p)), max(v, int(p))) for v in range(1, n) for p in [rng.integers(0, v)]}
    for a, b in rng.integers(0, n, size=(int(rng.integers(0, 2 * n + 1)), 2)).tolist():
        if a != b:
            edges.add((min(a, b), max(a, b)))
    e = np.array(sorted(edges), np.int32).reshape(-1, 2)
    w = rng.integers(1, int(rng.choice([1, 2, 3, 1000])) + 1, size=len(e)).astype(np.int32)
    g = gg.ArcGraph(n, np.concatenate([e[:, 0], e[:, 1]]), np.concatenate([e[:, 1], e[:, 0]]),
                    np.concatenate([w, w]))
    return g, rng.permutation(n).astype(np.int32)


@pytest.mark.parametrize("seed", range(24))
def test_random_small_graphs_every_form(seed):
    """Tree shapes the road graphs do not have (paths of bags, one-entry
    bags, wide top levels) and tie-heavy weights: full BC in the three forms
    within 1e-9 of the oracle, on the min-degree index and on a random
    contraction order, and the bag sweep's forward rows bit for bit."""
    g, order = _random_small_graph(seed)
    csr = oracle.CSR.from_arcgraph(g)
    ref, ovf = oracle.brandes(csr)
    assert not ovf
    G = bcx.Graph.from_arcs(g.n, g.u, g.v, g.w)
    for ix in (bcx.Index.build_mhl(G), bcx.Index.build_order(G, order)):
        ctx = bcx.Context(G, ix, 0)
        for mode in FORMS:
            bc, o = ctx.full_bc(mode)
            assert not o and close_bc(bc, ref), f"seed {seed} n {g.n}, mode {mode}"
        src = np.arange(g.n, dtype=np.int32)
        d, s, _, _ = ctx.query_rows(src, bcx.MODE_SWEEP)
        for i in range(g.n):
            od, osg, _, _ = oracle.sssp(csr, i)
            assert np.array_equal(d[i].astype(np.int64), od) and np.array_equal(s[i], osg), f"seed {seed} row {i}"
        ctx.close()


def test_ragged_and_empty_source_lists(grid):
    g, csr, ctx = grid
    ctx.set_tuning(sweep_batch=4096, batch=256)
    for k in (0, 1, 63, 65, 129):
        src = gg.permutation(17, k + 1, g.n)[:k].astype(np.int32)
        ref = oracle.brandes(csr, src)[0] if k else np.zeros(g.n)
        for mode in FORMS:
            out, ovf = ctx.bc_sources(src, mode)
            assert not ovf and close_bc(out, ref), f"{k} sources, mode {mode}"


def test_repeated_sources_count_twice(grid):
    g, csr, ctx = grid
    src = np.array([5, 5, 700], np.int32)
    ref = oracle.brandes(csr, np.array([5, 700], np.int32))[0] + oracle.brandes(csr, np.array([5], np.int32))[0]
    for mode in FORMS:
        out, _ = ctx.bc_sources(src, mode)
        assert close_bc(out, ref), f"mode {mode}"


# ---- QBC distance-only scan (H2H labels) ------------------------------------

@pytest.mark.parametrize("maker,order", [
    (lambda: gg.make("grid1k"), False),
    (lambda: gg.make("grid1k"), True),                      # order-file (QBC) tree
    (lambda: gg.grid_graph(k=20, seed=3, wmax=2), False),   # many ties
    (lambda: gg.road_graph(20_000, 25_000, seed=11), False),
])
def test_h2h_distances_match_dijkstra(maker, order):
    g = maker()
    G = bcx.Graph.from_arcs(g.n, g.u, g.v, g.w)
    ix = (bcx.Index.build_order(G, gg.permutation(5, 2, g.n).argsort().astype(np.int32))
          if order else bcx.Index.build_mhl(G))
    ix.build_h2h()
    ctx = bcx.Context(G, ix, 0)
    csr = oracle.CSR.from_arcgraph(g)
    src = gg.permutation(6, 1, g.n)[: min(g.n, 96)].astype(np.int32)
    got = ctx.h2h_distances(src)
    for i, s in enumerate(src.tolist()):
        assert np.array_equal(got[i].astype(np.int64), oracle.sssp(csr, s)[0]), f"source {s}"


def test_h2h_distances_in_several_chunks():
    # More than 65,535 sources: several launches, each with its own row reorder.
    g = gg.make("grid1k")
    G = bcx.Graph.from_arcs(g.n, g.u, g.v, g.w)
    ix = bcx.Index.build_mhl(G)
    ix.build_h2h()
    ctx = bcx.Context(G, ix, 0)
    csr = oracle.CSR.from_arcgraph(g)
    uniq = gg.permutation(8, 1, g.n)[:40].astype(np.int32)
    ref = np.stack([oracle.sssp(csr, int(s))[0] for s in uniq]).astype(np.int32)
    reps = 65_535 // len(uniq) + 2
    got = ctx.h2h_distances(np.tile(uniq, reps))
    assert got.shape == (reps * len(uniq), g.n)
    assert np.array_equal(got.reshape(reps, len(uniq), g.n), np.broadcast_to(ref, (reps, len(uniq), g.n)))


def test_h2h_distances_from_the_reference_h2h_index(golden_dir):
    """QBC's H2H index as the reference's host code builds it for the grid
    (distance rows, optimizeVS separators, its own P2H branch tables), handed
    over through bcx_index_h2h_from_arrays (the seam of moveDisIndexToDevice,
    QBC.cu:11-127): the device scan from every source equals Dijkstra, and the
    distances of bcx's own H2H build."""
    g = gg.make("grid1k")
    G = bcx.Graph.from_arcs(g.n, g.u, g.v, g.w)
    ref = dict(np.load(golden_dir / "grid1k_h2hfull.npz"))
    ix = bcx.Index.build_order(G, ref["order"])
    ix.h2h_from_arrays(ref)
    ctx = bcx.Context(G, ix, 0)
    src = np.arange(g.n, dtype=np.int32)
    got = ctx.h2h_distances(src)
    ctx.close()
    csr = oracle.CSR.from_arcgraph(g)
    for s in range(g.n):
        assert np.array_equal(got[s].astype(np.int64), oracle.sssp(csr, s)[0]), f"source {s}"
    own = bcx.Index.build_order(G, ref["order"])
    own.build_h2h()
    ctx = bcx.Context(G, own, 0)
    assert np.array_equal(ctx.h2h_distances(src), got)
    ctx.close()


def test_h2h_distances_need_labels(grid):
    g, csr, ctx = grid
    with pytest.raises(bcx.BcxError):
        ctx.h2h_distances(np.array([0], np.int32))


# ---- bag-sweep form ---------------------------------------------------------

@pytest.mark.parametrize("maker", [
    lambda: gg.make("grid1k"),
    lambda: gg.grid_graph(k=20, seed=3, wmax=2),   # many ties
    lambda: gg.unit_grid(12),                       # counts up to C(22, 11)
    lambda: gg.road_graph(3000, 3750, seed=21),
])
def test_sweep_full_bc(maker):
    g = maker()
    G, ix, ctx = build(g)
    ref, _ = oracle.brandes(oracle.CSR.from_arcgraph(g))
    for batch in (64, 4096):
        ctx.set_tuning(sweep_batch=batch)
        out, ovf = ctx.full_bc(bcx.MODE_SWEEP)
        assert not ovf
        assert close_bc(out, ref), f"batch {batch}: max rel err " \
            f"{np.max(np.abs(out - ref) / np.maximum(np.abs(ref), 1)):.3e}"


def test_sweep_two_kernel_forms_agree_bitwise(road):
    # More tiles in a launch than compute units selects the second form of the
    # tile kernel (two workgroups per unit, one piece per step); launches of
    # fewer tiles use the first.  Same tiles, same summation orders: the
    # results must be bitwise equal.
    g, csr, ctx = road
    default = ctx.tuning()["sweep_batch"]
    _, cus = ctx.sweep_slots()
    src = np.arange(g.n, dtype=np.int32)
    assert (g.n + 63) // 64 > cus, "the launch must hold more tiles than compute units"
    ctx.set_tuning(sweep_batch=default)
    two, o1 = ctx.bc_sources(src, bcx.MODE_SWEEP)
    ctx.set_tuning(sweep_batch=64 * max(1, cus // 2))
    one, o2 = ctx.bc_sources(src, bcx.MODE_SWEEP)
    ctx.set_tuning(sweep_batch=default)
    assert not o1 and not o2
    assert np.array_equal(two.view(np.uint64), one.view(np.uint64))


def test_sweep_pooled_slots_bitwise(road):
    """A launch of more tiles than state slots: the one-workgroup form keeps
    one slot per compute unit and the launch's workgroups take free slots as
    others finish (a slot's records reused by several tiles in one launch).
    Same tiles, same summation orders: bitwise equal to launches of at most
    one tile per slot, and the oracle's BC."""
    g, csr, ctx = road
    default = ctx.tuning()["sweep_batch"]
    _, cus = ctx.sweep_slots()
    tiles = (g.n + 63) // 64
    assert tiles > cus, "needs more tiles than compute units"
    ctx.set_tuning(sweep_batch=64 * tiles, sweep_form=bcx.SWEEP_FORM_ONE_PER_CU)
    ctx.reset_timing()
    pooled, o1 = ctx.full_bc(bcx.MODE_SWEEP)
    assert ctx.timing()["scan_launches"] == 1
    slots, _ = ctx.sweep_slots()
    assert slots == cus, f"{slots} slots for {tiles} tiles"
    ctx.set_tuning(sweep_batch=64 * cus)
    ctx.reset_timing()
    split, o2 = ctx.full_bc(bcx.MODE_SWEEP)
    assert ctx.timing()["scan_launches"] == -(-tiles // cus)
    ctx.set_tuning(sweep_batch=default, sweep_form=bcx.SWEEP_FORM_AUTO)
    assert not o1 and not o2
    assert np.array_equal(pooled.view(np.uint64), split.view(np.uint64))
    ref, _ = oracle.brandes(csr, threads=oracle.default_threads())
    assert close_bc(pooled, ref)


def test_sweep_repeated_job_keeps_its_plan(road):
    """A job over the same sources as the last one keeps the uploaded plan (a
    repeated full BC); another job in between plans anew.  Every full BC is
    bitwise the first."""
    g, csr, ctx = road
    first, o1 = ctx.full_bc(bcx.MODE_SWEEP)
    again, o2 = ctx.full_bc(bcx.MODE_SWEEP)
    src = np.arange(0, g.n, 7, dtype=np.int32)
    part, o3 = ctx.bc_sources(src, bcx.MODE_SWEEP)
    last, o4 = ctx.full_bc(bcx.MODE_SWEEP)
    assert not (o1 or o2 or o3 or o4)
    assert np.array_equal(again.view(np.uint64), first.view(np.uint64))
    assert np.array_equal(last.view(np.uint64), first.view(np.uint64))
    want, _ = oracle.brandes(csr, src, threads=oracle.default_threads())
    assert close_bc(part, want)


def test_sweep_scattered_sources(road):
    g, csr, ctx = road
    src = gg.permutation(6, 2, g.n)[:700].astype(np.int32)  # not tree-consecutive
    ctx.set_tuning(sweep_batch=256)
    out, ovf = ctx.bc_sources(src, bcx.MODE_SWEEP)
    ref, _ = oracle.brandes(csr, src)
    assert not ovf
    assert close_bc(out, ref)


def test_sweep_overflow_flag():
    g = gg.unit_grid(35)
    G, ix, ctx = build(g)
    _, ovf = ctx.bc_sources(np.array([0], np.int32), bcx.MODE_SWEEP)
    assert ovf


# ---- distance / count rows of every form, bit-exact -------------------------

def _rows_match(g, csr, got, sources, hop=True):
    d, s, h, _ = got
    for i, src in enumerate(np.asarray(sources).tolist()):
        od, osg, ohp, _ = oracle.sssp(csr, src)
        assert np.array_equal(d[i].astype(np.int64), od), f"distances from {src}"
        assert np.array_equal(s[i], osg), f"counts from {src}"
        if hop and h is not None:
            assert np.array_equal(h[i], ohp), f"layers from {src}"


SWEEP_FORMS = [bcx.SWEEP_FORM_ONE_PER_CU, bcx.SWEEP_FORM_TWO_PER_CU]


def test_rows_every_form_grid_all_sources(grid):
    """The label-scan tile kernel and the bag sweep's forward pass (both
    tile-kernel forms) give the oracle's distances and counts from every
    source of the grid, bit for bit (and the label scan its layers)."""
    g, csr, ctx = grid
    src = np.arange(g.n, dtype=np.int32)
    ctx.set_tuning(batch=256)
    _rows_match(g, csr, ctx.query_rows(src, bcx.MODE_BATCHED), src)
    for form in SWEEP_FORMS:
        ctx.set_tuning(sweep_form=form)
        _rows_match(g, csr, ctx.query_rows(src[::-1].copy(), bcx.MODE_SWEEP), src[::-1])
    ctx.set_tuning(sweep_form=bcx.SWEEP_FORM_AUTO, batch=4096)


def test_rows_sweep_needs_no_hop(grid):
    g, csr, ctx = grid
    d = np.zeros(g.n, np.int32)
    s = np.zeros(g.n, np.uint64)
    h = np.zeros(g.n, np.uint32)
    rc = bcx.lib().bcx_query_dist_count_mode(ctx._h, bcx.MODE_SWEEP, np.array([0], np.int32), 1, d, s,
                                             h.ctypes.data)
    assert rc == -1


@pytest.mark.parametrize("k", [34, 35])
def test_rows_unit_grid_saturation_every_form(k):
    """C(2k-2, k-1) paths corner to corner: fits uint64 at k = 34, saturates
    (flagged) at k = 35, identically in every form and in the oracle."""
    from math import comb

    g = gg.unit_grid(k)
    G, ix, ctx = build(g)
    csr = oracle.CSR.from_arcgraph(g)
    src = np.array([0, k - 1, k * k - 1, k * (k // 2) + k // 2], np.int32)
    od, osg, _, oovf = oracle.sssp(csr, 0)
    assert oovf == (k == 35)
    want = comb(2 * k - 2, k - 1)
    runs = [(bcx.MODE_BATCHED, None)] + [(bcx.MODE_SWEEP, f) for f in SWEEP_FORMS]
    for mode, form in runs:
        if form is not None:
            ctx.set_tuning(sweep_form=form)
        d, s, h, ovf = ctx.query_rows(src, mode)
        assert ovf == (k == 35), f"mode {mode} form {form}"
        assert int(s[0, k * k - 1]) == (want if k == 34 else 2**64 - 1)
        _rows_match(g, csr, (d, s, h, ovf), src)


def test_rows_wide_distances_small_count_field():
    """The sweep keeps a lane's path count in its distance word when it fits
    above the distance bits, else in the count slot.  A 34 x 34 unit grid with
    every weight 7,500,000 needs 30 distance bits (twice the farthest root
    distance, 37 hops, is 555,000,000 >= 2^29), leaving counts 1 to 3 in the
    word and the binomial counts up to C(66, 33) in the slot: rows in both
    forms, and the label scan's, equal the oracle's bit for bit, and so does
    full BC."""
    from math import comb

    k, wgt = 34, 7_500_000
    u = gg.unit_grid(k)
    g = gg.ArcGraph(u.n, u.u, u.v, (u.w.astype(np.int64) * wgt).astype(np.int32))
    G, ix, ctx = build(g)
    csr = oracle.CSR.from_arcgraph(g)
    src = np.array([0, k - 1, k * k - 1, k * (k // 2) + k // 2, 5 * k + 3], np.int32)
    runs = [(bcx.MODE_BATCHED, None)] + [(bcx.MODE_SWEEP, f) for f in SWEEP_FORMS]
    for mode, form in runs:
        if form is not None:
            ctx.set_tuning(sweep_form=form)
        d, s, h, ovf = ctx.query_rows(src, mode)
        assert not ovf, f"mode {mode} form {form}"
        assert int(d[0, k * k - 1]) == 2 * (k - 1) * wgt
        assert int(s[0, k * k - 1]) == comb(2 * k - 2, k - 1)
        _rows_match(g, csr, (d, s, h, ovf), src)
    ref, _ = oracle.brandes(csr)
    for form in SWEEP_FORMS:
        ctx.set_tuning(sweep_form=form)
        bc, ovf = ctx.full_bc(bcx.MODE_SWEEP)
        assert not ovf and close_bc(bc, ref), f"form {form}"
    ctx.close()


def test_rows_road_every_form(road):
    g, csr, ctx = road
    src = np.concatenate([np.arange(200, 264), gg.permutation(31, 1, g.n)[:16]]).astype(np.int32)
    ctx.set_tuning(batch=64)
    _rows_match(g, csr, ctx.query_rows(src, bcx.MODE_BATCHED), src)
    for form in SWEEP_FORMS:
        ctx.set_tuning(sweep_form=form)
        _rows_match(g, csr, ctx.query_rows(src, bcx.MODE_SWEEP), src)
    ctx.set_tuning(sweep_form=bcx.SWEEP_FORM_AUTO, batch=4096)


def test_rows_la_golden_pairs_every_form():
    """On the reference's LosAngeles graph, the rows of the label-scan tile
    kernel and of the bag sweep's forward pass equal the answers of the
    reference's own index code on its golden pairs."""
    golden = ROOT / "tests" / "golden"
    la = np.load(golden / "la_graph.npz")
    ref = np.load(golden / "la_tc.npz")
    n = int(la["n"])
    G = bcx.Graph.from_arcs(n, la["u"], la["v"], la["w"])
    ix = bcx.Index.build_mhl(G)
    ctx = bcx.Context(G, ix, 0)
    q = ref["queries"].astype(np.int64)
    srcs = np.unique(q[:, 0])
    src0 = (srcs - 1).astype(np.int32)
    runs = [(bcx.MODE_BATCHED, None)] + [(bcx.MODE_SWEEP, f) for f in SWEEP_FORMS]
    for mode, form in runs:
        if form is not None:
            ctx.set_tuning(sweep_form=form)
        d, s, h, _ = ctx.query_rows(src0, mode)
        for i, sv in enumerate(srcs.tolist()):
            rows = q[q[:, 0] == sv]
            t = rows[:, 1] - 1
            assert np.array_equal(d[i, t].astype(np.int64), rows[:, 2]), f"distances from {sv}, mode {mode}"
            assert np.array_equal(s[i, t].astype(np.int64), rows[:, 3]), f"counts from {sv}, mode {mode}"
            if h is not None:
                assert np.array_equal(h[i, t].astype(np.int64), rows[:, 4]), f"layers from {sv}"
    ctx.close()


# ---- unpack, memory-capped launches, captured jobs ------------------------------

@pytest.mark.parametrize("maker", [
    lambda: gg.make("grid1k"),
    lambda: gg.road_graph(3000, 3750, seed=21),
])
def test_sweep_wide_unpack_path(maker):
    """Every depth level through the grid-wide unpack launches (threshold 1)
    against the oracle, and against the default split within rounding."""
    g = maker()
    G, ix, ctx = build(g)
    ref, _ = oracle.brandes(oracle.CSR.from_arcgraph(g), threads=oracle.default_threads())
    base, _ = ctx.full_bc(bcx.MODE_SWEEP)
    ctx.set_tuning(unpack_wide=1)
    wide, ovf = ctx.full_bc(bcx.MODE_SWEEP)
    assert not ovf
    assert close_bc(wide, ref)
    assert np.all(np.abs(wide - base) <= 1e-12 * np.maximum(np.abs(base), 1.0))


def test_sweep_memory_capped_launches(road):
    """Launches cut short by the memory cap, the label-scan workspace released
    first: a cap that holds one state slot per compute unit but not two runs
    the job as one pooled launch in the one-workgroup form; a cap below one
    slot per unit runs it as several launches of one tile per slot.  Same
    tiles, same summation orders: both bitwise equal to the uncapped job, and
    the oracle's BC."""
    g, csr, _ = road
    G, ix, ctx = build(g)  # library default tuning (the shared context's is changed by other tests)
    _, cus = ctx.sweep_slots()
    tiles = (g.n + 63) // 64
    if tiles <= cus + 8:
        pytest.skip("needs more tiles than compute units")
    ctx.set_tuning(batch=256)
    ctx.bc_sources(np.arange(256, dtype=np.int32), bcx.MODE_BATCHED)  # holds the workspace
    per_slot = (g.n + 1) * 768
    ctx.reset_timing()
    whole, o0 = ctx.full_bc(bcx.MODE_SWEEP)
    assert ctx.timing()["scan_launches"] == 1
    ctx.set_tuning(sweep_mem_limit=int(1.5 * cus * per_slot))  # one slot per unit, not two
    ctx.bc_sources(np.arange(256, dtype=np.int32), bcx.MODE_BATCHED)  # holds the workspace again
    ctx.reset_timing()
    pooled, o1 = ctx.full_bc(bcx.MODE_SWEEP)
    assert ctx.timing()["scan_launches"] == 1
    assert ctx.sweep_slots()[0] == cus
    want = cus // 2
    ctx.set_tuning(sweep_mem_limit=want * per_slot)  # fewer slots than units
    ctx.reset_timing()
    split, o2 = ctx.full_bc(bcx.MODE_SWEEP)
    launches = ctx.timing()["scan_launches"]
    slots, _ = ctx.sweep_slots()
    ctx.set_tuning(sweep_mem_limit=0)
    assert 1 <= slots < want and launches == -(-tiles // slots), f"{launches} launches, {slots} slots"
    assert not o0 and not o1 and not o2
    assert np.array_equal(pooled.view(np.uint64), whole.view(np.uint64))
    assert np.array_equal(split.view(np.uint64), whole.view(np.uint64))
    ref, _ = oracle.brandes(csr, threads=oracle.default_threads())
    assert close_bc(whole, ref)
    ctx.close()


@pytest.mark.parametrize("pooled", [False, True])
def test_sweep_captured_job_replays(road, pooled):
    """The captured launch sequence of a job, replayed several times with
    eager (profiled, uncaptured) jobs in between, gives the eager result bit
    for bit every time.  pooled: the one-workgroup form's 256 slots under
    launches of 300 tiles (the slot flags cleared by a kernel in the graph)."""
    g, csr, ctx = road
    saved = ctx.tuning()
    if pooled:
        src = np.arange(g.n, dtype=np.int32)
        ctx.set_tuning(sweep_batch=300 * 64, sweep_form=bcx.SWEEP_FORM_ONE_PER_CU)
        assert ctx.sweep_capacity() == 300 * 64
        slots, cus = ctx.sweep_slots()
        assert slots == cus < 300, f"{slots} slots"
    else:
        src = np.arange(0, g.n, 3, dtype=np.int32)
    try:
        ctx.set_profiling(True)
        eager, _ = ctx.bc_sources(src, bcx.MODE_SWEEP)
        ctx.set_profiling(False)
        for rep in range(4):
            got, _ = ctx.bc_sources(src, bcx.MODE_SWEEP)
            assert np.array_equal(got.view(np.uint64), eager.view(np.uint64)), f"replay {rep}"
            ctx.set_profiling(True)
            other, _ = ctx.bc_sources(src[:777], bcx.MODE_SWEEP)
            ctx.set_profiling(False)
    finally:
        ctx.set_profiling(False)
        ctx.set_tuning(sweep_batch=saved["sweep_batch"], sweep_form=saved["sweep_form"])
    ref, _ = oracle.brandes(csr, src, threads=oracle.default_threads())
    assert close_bc(eager, ref)


# ---- RQ admission ----------------------------------------------------------------

@pytest.mark.parametrize("mode", [bcx.MODE_SWEEP, bcx.MODE_BATCHED])
def test_rq_admission_busy_readmit_cancel(grid, mode):
    """A budget that holds one request at a time: the second is refused with
    BCX_EBUSY, admitted once the first completes; a request cancelled in the
    middle of its cycle returns its bytes; results match the oracle and the
    request alone, bit for bit."""
    g, csr, ctx = grid
    ref, _ = oracle.brandes(csr)
    saved = ctx.tuning()
    ctx.set_tuning(batch=128, sweep_batch=128)
    a = np.array([3, 9, 400, 1000], np.int32)
    b = np.arange(0, 1024, 5, dtype=np.int32)
    probe = bcx.RQServer(ctx, 1 << 30, mode)
    fixed, _ = probe.memory()
    need_b = probe.bytes_needed(len(b))
    probe.close()
    srv = bcx.RQServer(ctx, fixed + need_b, mode)  # room for b alone, or a alone
    with pytest.raises(bcx.BcxError):
        bcx.RQServer(ctx, fixed - 1, mode)
    ra = srv.submit(a)
    assert ra is not None
    assert srv.submit(b) is None, "b must not fit while a is in flight"
    res_a = None
    steps = 0
    while res_a is None:
        srv.step()
        steps += 1
        res_a = srv.poll(ra)
    assert srv.memory()[0] == fixed
    rb = srv.submit(b)
    assert rb is not None
    srv.step()
    srv.step()
    srv.cancel(rb)
    assert srv.memory()[0] == fixed
    rb = srv.submit(b)
    res_b = None
    while res_b is None:
        srv.step()
        res_b = srv.poll(rb)
    assert close_bc(res_a, ref[a])
    assert close_bc(res_b, ref[b])
    alone = bcx.RQServer(ctx, 1 << 30, mode)
    rid = alone.submit(a)
    got = None
    while got is None:
        alone.step()
        got = alone.poll(rid)
    assert np.array_equal(got.view(np.uint64), res_a.view(np.uint64))
    ctx.set_tuning(**saved)


# ---- persisted index and the host-array seam ----------------------------------

def test_bc_from_loaded_image_and_cli_restart(tmp_path, grid):
    """A saved image (with H2H labels) loaded in a new context gives the
    oracle's BC in both forms and the same H2H distances; MHLQBC restarted
    from the image (BCX_INDEX_LOAD) writes the same BC."""
    g, csr, _ = grid
    ref, _ = oracle.brandes(csr)
    G = bcx.Graph.from_arcs(g.n, g.u, g.v, g.w)
    ix = bcx.Index.build_mhl(G)
    ix.build_h2h()
    p = tmp_path / "grid.idx"
    ix.save(p)
    ctx = bcx.Context(G, bcx.Index.load(p), 0)
    for mode in (bcx.MODE_SWEEP, bcx.MODE_BATCHED):
        bc, ovf = ctx.full_bc(mode)
        assert not ovf and close_bc(bc, ref), f"mode {mode}"
    d = ctx.h2h_distances(np.array([0, 500], np.int32))
    assert np.array_equal(d[1].astype(np.int64), oracle.sssp(csr, 500)[0])
    ctx.close()
    gp = tmp_path / "g.txt"
    gg.write_graph(g, gp)
    exe = ROOT / "gpu_betweennesscentrality_amd" / "bin" / "MHLQBC"
    env = dict(os.environ, BCX_INDEX_LOAD=str(p), BCX_BC_DUMP=str(tmp_path / "bc.bin"))
    out = subprocess.run([exe, gp, "0"], env=env, capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    assert close_bc(np.fromfile(tmp_path / "bc.bin", np.float64), ref)
    other = gg.grid_graph(k=32, seed=2)
    op = tmp_path / "other.txt"
    gg.write_graph(other, op)
    out = subprocess.run([exe, op, "0"], env=env, capture_output=True, text=True, timeout=120)
    assert out.returncode == 1 and "another graph" in out.stderr


def test_bc_from_reference_treedecomp_through_the_seam():
    """The reference's own host index for the grid (its label rows and bags,
    dumped by the reference's TreeDecomp code) handed over through
    bcx_index_from_arrays: BC in both forms equals the oracle's."""
    golden = ROOT / "tests" / "golden"
    full = dict(np.load(golden / "grid1k_tcfull.npz"))
    tc = np.load(golden / "grid1k_tc.npz")
    n = int(tc["n"])
    G = bcx.Graph.from_arcs(n, tc["u"], tc["v"], tc["w"])
    ctx = bcx.Context(G, bcx.Index.from_arrays(full, G), 0)
    ref, _ = oracle.brandes(oracle.CSR(n, tc["u"], tc["v"], tc["w"]))
    for mode in (bcx.MODE_SWEEP, bcx.MODE_BATCHED):
        bc, ovf = ctx.full_bc(mode)
        assert not ovf and close_bc(bc, ref), f"mode {mode}"
    ctx.close()


def test_weights_summing_past_inf():
    """Graphs the old total-weight cap refused: a star whose weights sum past
    the label INF, and a triangle with an arc longer than any distance."""
    n = 3001
    e = np.array([(0, i) for i in range(1, n)], np.int32)
    star = gg._both_ways(n, e[:, 0], e[:, 1], np.full(n - 1, 400_000, np.int64))
    G, ix, ctx = build(star)
    want = np.zeros(n)
    want[0] = (n - 1) * (n - 2) / 2
    for mode in FORMS:
        bc, ovf = ctx.full_bc(mode)
        assert not ovf and close_bc(bc, want), f"star, mode {mode}"
    tri = gg._both_ways(3, np.array([0, 1, 0]), np.array([1, 2, 2]), np.array([5, 5, 2_000_000_000]))
    G, ix, ctx = build(tri)
    for mode in FORMS:
        bc, ovf = ctx.full_bc(mode)
        assert not ovf and close_bc(bc, np.array([0.0, 1.0, 0.0])), f"triangle, mode {mode}"


def test_autotune_picks_a_sweep_launch_and_stays_exact(road):
    """The autotuner times bag-sweep launch sizes and tile-kernel forms; the
    tuning it keeps gives the oracle's BC."""
    g, csr, _ = road
    G, ix, ctx = build(g)
    _, cus = ctx.sweep_slots()
    default = ctx.tuning()["sweep_batch"]
    best = ctx.autotune(gg.permutation(12, 1, g.n)[:4096].astype(np.int32))
    assert best["sweep_form"] in (bcx.SWEEP_FORM_AUTO, bcx.SWEEP_FORM_ONE_PER_CU, bcx.SWEEP_FORM_TWO_PER_CU)
    assert best["sweep_batch"] == default or (
        best["sweep_batch"] % (64 * cus) == 0 and best["sweep_batch"] // (64 * cus) in (2, 4, 8, 16, 32))
    assert ctx.tuning() == best
    src = gg.permutation(13, 1, g.n)[:3000].astype(np.int32)
    got, ovf = ctx.bc_sources(src, bcx.MODE_SWEEP)
    want, _ = oracle.brandes(csr, src, threads=oracle.default_threads())
    assert not ovf and close_bc(got, want)
    ctx.close()


# ---- label rows computed on the device -----------------------------------------

def _la_graph():
    la = np.load(ROOT / "tests" / "golden" / "la_graph.npz")
    return la, bcx.Graph.from_arcs(int(la["n"]), la["u"], la["v"], la["w"])


@pytest.mark.parametrize("case", ["grid1k", "ties", "unit35", "road3k", "la", "la_order"])
def test_device_label_rows_equal_host_rows(case):
    """An index that leaves its label rows to the device (BCX_INDEX_DEVICE_LABELS):
    the rows the context computes equal the host build's bit for bit
    (distance, saturating uint64 count, max hop), on a tie-heavy grid, a
    unit grid whose counts saturate, a road graph and LosAngeles with the
    min-degree and the shipped order; BC is bitwise the same in the bag-sweep
    and label-scan forms."""
    order = None
    if case in ("la", "la_order"):
        la, G = _la_graph()
        order = la["order"] if case == "la_order" else None
    else:
        g = {"grid1k": lambda: gg.make("grid1k"), "ties": lambda: gg.grid_graph(k=20, seed=3, wmax=2),
             "unit35": lambda: gg.unit_grid(35), "road3k": lambda: gg.road_graph(3000, 3750, seed=21)}[case]()
        G = bcx.Graph.from_arcs(g.n, g.u, g.v, g.w)
    if order is None:
        host, dev = bcx.Index.build_mhl(G), bcx.Index.build_mhl(G, device_labels=True)
    else:
        host, dev = bcx.Index.build_order(G, order), bcx.Index.build_order(G, order, device_labels=True)
    want = host.export()
    ch, cd = bcx.Context(G, host, 0), bcx.Context(G, dev, 0)
    got = cd.export_labels()
    for k in ("label_off", "dis", "cnt", "hop"):
        assert np.array_equal(got[k], want[k]), f"{case}: device rows differ in {k}"
    assert np.array_equal(ch.export_labels()["cnt"], want["cnt"])  # the export path itself
    for mode in (bcx.MODE_SWEEP, bcx.MODE_BATCHED):
        a, oa = ch.full_bc(mode)
        b, ob = cd.full_bc(mode)
        assert oa == ob and np.array_equal(a.view(np.uint64), b.view(np.uint64)), f"{case}, mode {mode}"
    ch.close()
    cd.close()


def test_device_label_rows_distance_bound():
    """The device build refuses graphs whose distances may reach the label INF,
    as the host build does, and accepts those the host build accepts."""
    n = 3001
    e = np.array([(0, i) for i in range(1, n)], np.int32)
    star = gg._both_ways(n, e[:, 0], e[:, 1], np.full(n - 1, 400_000, np.int64))
    G = bcx.Graph.from_arcs(star.n, star.u, star.v, star.w)
    ctx = bcx.Context(G, bcx.Index.build_mhl(G, device_labels=True), 0)
    bc, ovf = ctx.full_bc(bcx.MODE_SWEEP)
    assert not ovf and bc[0] == (n - 1) * (n - 2) / 2
    ctx.close()
    far = gg._both_ways(3, np.array([0, 1]), np.array([1, 2]), np.array([600_000_000, 600_000_000]))
    G = bcx.Graph.from_arcs(far.n, far.u, far.v, far.w)
    with pytest.raises(bcx.BcxError):
        bcx.Index.build_mhl(G)
    with pytest.raises(bcx.BcxError):
        bcx.Context(G, bcx.Index.build_mhl(G, device_labels=True), 0)
