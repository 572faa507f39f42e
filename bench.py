"""Benchmark of the betweenness hot path on one MI355X.

A step is one batched launch sequence over one batch of source vertices on
the seeded road-like 320,000-vertex graph with its MHL index resident in
HBM: the bag-sweep form (3 tiles of 64 sources per compute unit, 49,152 on
MI355X: per-tile sweep kernel, tile
reduction; the shortcut unpacking runs once per full BC and is amortized).
The value is whole-job full-BC throughput in sources/s; the label-scan form,
the plain one-source-per-launch form, the RQ serving latencies and the CPU
baseline are reported beside it.

  python bench.py [--steps K] [--warmup W] [--workload road320k] [--gpus 1]
  python bench.py --profile-only            # one batch of each form + a 1 GiB copy
  python bench.py --summarize-trace CSV     # rocprofv3 kernel_stats.csv -> markdown
  python bench.py --summarize-pmc FETCH_CSV WRITE_CSV   # HBM traffic per launch
  python bench.py --summarize-counters CSV...  # any counters of the sweep kernel
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
SWEEP_KERNEL = "sweep_tile_kernel"
COPY_KERNEL = "__amd_rocclr_copyBuffer"
COPY_BYTES = 1 << 30  # the --profile-only device copy
XCDS, CUS, SIMDS = 8, 256, 1024  # MI355X: accelerator dies, compute units, SIMDs
PROFILES = ROOT / "profiles"


def sweep_algorithmic_bytes(n: int, bags: int, inner: int, height: int, sources: int) -> float:
    """Bytes one sweep launch must move for `sources` sources, a lower bound.

    Per source: both passes read one 4-byte distance word per bag entry
    (8 * bags); the forward pass writes one distance word per node (4 * n);
    the backward pass, which runs only on the `inner` nodes (nodes with
    descendants), reads each one's distance word back and writes its 8-byte
    quotient (12 * inner).  Every inner node that is not an ancestor of the
    source (at most `height` are) reaches its distance through at least one
    tight bag entry, whose ancestor's backward step reads the node's quotient
    (8 * (inner - height)).  Count words (written and read only where a path
    count is not 1) and the quotients of further tight entries are
    data-dependent and not counted.
    """
    return float(sources) * (8.0 * bags + 4.0 * n + 12.0 * inner + 8.0 * max(0, inner - height))


def _kernel_rows(csv_path: str, counter: str):
    import csv
    rows = {}
    for r in csv.DictReader(open(csv_path)):
        if r.get("Counter_Name") != counter:
            continue
        name = r["Kernel_Name"]
        rows.setdefault(name, []).append(float(r["Counter_Value"]))
    return rows


def summarize_pmc(fetch_csv: str, write_csv: str) -> str:
    """HBM traffic of the sweep kernel from FETCH_SIZE and WRITE_SIZE runs.

    Each counter is scaled by what it reads for the 1 GiB device copy of the
    same run (a copy fetches and writes exactly 1 GiB).
    """
    f = _kernel_rows(fetch_csv, "FETCH_SIZE")
    w = _kernel_rows(write_csv, "WRITE_SIZE")

    def pick(rows, sub):
        for k, v in rows.items():
            if sub in k:
                return max(v)
        raise SystemExit(f"no {sub} row")

    fk, wk = pick(f, SWEEP_KERNEL), pick(w, SWEEP_KERNEL)
    fc, wc = pick(f, COPY_KERNEL), pick(w, COPY_KERNEL)
    ff = (COPY_BYTES / 1024.0) / fc
    wf = (COPY_BYTES / 1024.0) / wc
    fetch = fk * 1024.0 * ff
    write = wk * 1024.0 * wf
    return (f"| counter | sweep kernel (KiB) | 1 GiB copy (KiB) | scale | bytes per launch |\n"
            f"|---|---|---|---|---|\n"
            f"| FETCH_SIZE | {fk:.0f} | {fc:.0f} | {ff:.3f} | {fetch:.4e} |\n"
            f"| WRITE_SIZE | {wk:.0f} | {wc:.0f} | {wf:.3f} | {write:.4e} |\n\n"
            f"traffic_bytes_per_launch: {fetch + write:.6e}\n")


def summarize_counters(csv_paths) -> str:
    """Every counter of the sweep kernel's dispatch in rocprofv3
    counter_collection.csv files (summed over the rows of that dispatch), and
    the shares they imply: per-SIMD issue busy fractions from the SQ counters
    (counted in 4-cycle units) and per-CU busy fractions of the texture
    address (TA) and data (TD) units, all over the kernel's cycles
    (GRBM_GUI_ACTIVE, one count per XCD)."""
    import csv
    tot = {}
    for path in csv_paths:
        disp = {}
        for r in csv.DictReader(open(path)):
            if SWEEP_KERNEL not in r["Kernel_Name"]:
                continue
            key = (r["Dispatch_Id"], r["Counter_Name"])
            disp[key] = disp.get(key, 0.0) + float(r["Counter_Value"])
        last = max(d for d, _ in disp) if disp else None
        for (d, name), v in disp.items():
            if d == last:
                tot[name] = v
    out = ["| counter | value |", "|---|---|"]
    out += [f"| {k} | {v:.4e} |" for k, v in sorted(tot.items())]
    cyc = tot.get("GRBM_GUI_ACTIVE", 0.0) / XCDS
    if cyc:
        out += ["", "| derived | value |", "|---|---|", f"| kernel cycles | {cyc:.4e} |"]
        for k, label in (("SQ_ACTIVE_INST_VALU", "VALU busy per SIMD"),
                         ("SQ_ACTIVE_INST_SCA", "scalar busy per SIMD")):
            if k in tot:
                out.append(f"| {label} | {4 * tot[k] / SIMDS / cyc:.1%} |")
        for k, label in (("TA_TA_BUSY_sum", "TA busy per CU"), ("TD_TD_BUSY_sum", "TD busy per CU"),
                         ("TD_TC_STALL_sum", "TD stalled on the L1 per CU")):
            if k in tot:
                out.append(f"| {label} | {tot[k] / CUS / cyc:.1%} |")
        if "SQ_WAVE_CYCLES" in tot:
            for k, label in (("SQ_WAIT_ANY", "wave cycles waiting on memory counters"),
                             ("SQ_WAIT_INST_ANY", "wave cycles waiting to issue")):
                if k in tot:
                    out.append(f"| {label} | {tot[k] / tot['SQ_WAVE_CYCLES']:.1%} |")
    return "\n".join(out) + "\n"


def summarize_trace(csv_path: str) -> str:
    import csv
    out = ["| kernel | calls | avg (us) | min (us) | max (us) | share |", "|---|---|---|---|---|---|"]
    for r in csv.DictReader(open(csv_path)):
        name = r["Name"].replace("(anonymous namespace)::", "").split("(")[0].replace("void ", "")
        out.append(f"| {name} | {r['Calls']} | {float(r['AverageNs']) / 1e3:.1f} | "
                   f"{float(r['MinNs']) / 1e3:.1f} | {float(r['MaxNs']) / 1e3:.1f} | "
                   f"{float(r['Percentage']):.2f}% |")
    return "\n".join(out) + "\n"


def committed_traffic(sources: int, workload: str):
    """Traffic per launch from the committed PMC summary of this workload
    (profiles/sweep_tile_<workload>_pmc.md), if it covers this launch size
    (summary lines `workload:` and `sources:`)."""
    path = PROFILES / f"sweep_tile_{workload}_pmc.md"
    if not path.exists():
        return None
    meta = {}
    for line in path.read_text().splitlines():
        if ":" in line and not line.startswith("|"):
            k, _, v = line.partition(":")
            meta[k.strip()] = v.strip()
    try:
        if meta.get("workload") != workload or int(meta.get("sources", -1)) != sources:
            return None
        return float(meta["traffic_bytes_per_launch"])
    except (KeyError, ValueError):
        return None


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=6)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="road320k")
    ap.add_argument("--batch", type=int, default=4096, help="label-scan sources per batch")
    ap.add_argument("--sweep-batch", type=int, default=0,
                    help="bag-sweep sources per launch (0: the library default, 3 tiles of 64 per CU)")
    ap.add_argument("--plain-sources", type=int, default=64)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-rq", action="store_true")
    ap.add_argument("--no-h2h", action="store_true", help="skip the QBC distance-only (H2H) scan")
    ap.add_argument("--rq-requests", type=int, default=6)
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--profile-only", action="store_true",
                    help="one label-scan batch, one sweep batch and a 1 GiB copy (for rocprofv3)")
    ap.add_argument("--summarize-trace", metavar="CSV",
                    help="print a markdown table of a rocprofv3 kernel_stats.csv and exit")
    ap.add_argument("--summarize-pmc", nargs=2, metavar=("FETCH_CSV", "WRITE_CSV"),
                    help="print the sweep kernel's HBM traffic from two counter_collection.csv")
    ap.add_argument("--summarize-counters", nargs="+", metavar="CSV",
                    help="print every counter of the sweep kernel from counter_collection.csv files")
    args = ap.parse_args()
    if args.summarize_counters:
        print(summarize_counters(args.summarize_counters), end="")
        return 0
    if args.summarize_trace:
        print(summarize_trace(args.summarize_trace), end="")
        return 0
    if args.summarize_pmc:
        print(summarize_pmc(*args.summarize_pmc), end="")
        return 0
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported "
              "(a full BC would shard by source ranges, one GPU each)", file=sys.stderr)
        return 2

    import gpu_betweennesscentrality_amd as bcx
    from gpu_betweennesscentrality_amd import graphgen as gg

    t0 = time.time()
    g = gg.make(args.workload)
    G = bcx.Graph.from_arcs(g.n, g.u, g.v, g.w)
    log(f"[bench] graph {args.workload}: n={g.n} arcs={g.arcs} ({time.time() - t0:.1f}s)")
    ix = bcx.Index.build_mhl(G)
    st = ix.stats()
    lbar = ix.mean_scan(1 << 18, 1)
    log(f"[bench] index: height={st['tree_height']} width={st['tree_width']} "
        f"entries={st['label_entries']} Lbar={lbar:.1f} build={st['build_seconds']:.1f}s")
    ctx = bcx.Context(G, ix, 0)
    n = g.n
    B = min(args.batch, n)
    ctx.set_tuning(batch=B)

    # Batches are blocks of the tree order (how a full sweep visits sources),
    # spread evenly over it so the sample covers shallow and deep subtrees.
    tree_perm = ix.tree_order()
    sweep_default = ctx.tuning()["sweep_batch"]

    def tree_batch(k: int, nb: int, size: int = B) -> np.ndarray:
        size = min(size, n)
        start = (k * (n - size)) // max(1, nb - 1) if nb > 1 else 0
        start -= start % 64
        return tree_perm[start:start + size]

    total = args.warmup + args.steps
    if args.profile_only:
        ctx.set_tuning(batch=B)
        ctx.bc_sources(tree_batch(1, 3, B), bcx.MODE_BATCHED)
        Bs = min(args.sweep_batch or sweep_default, n)
        ctx.set_tuning(sweep_batch=Bs)
        ctx.bc_sources(tree_batch(1, 3, Bs), bcx.MODE_SWEEP)
        ms = ctx.copy_probe(1 << 30)
        log(f"[bench] profile-only done; 1 GiB device copy {ms:.3f} ms")
        return 0

    def time_mode(mode: int, label: str, size: int):
        # One call per phase, as a full BC runs: the call launches one batch
        # per step and finishes the job once (BC accumulation, and for the
        # bag-sweep form the shortcut unpacking).
        ctx.set_tuning(**({"sweep_batch": size} if mode == bcx.MODE_SWEEP else {"batch": size}))

        def job(k0: int, k1: int) -> np.ndarray:
            # Batches k0..k1-1; where they would cover the graph, every
            # source once (a whole full BC).
            if (k1 - k0) * size >= n:
                return tree_perm
            return np.concatenate([tree_batch(k, total, size) for k in range(k0, k1)])

        if args.warmup:
            ctx.bc_sources(job(0, args.warmup), mode)
        src = job(args.warmup, total)
        t = time.perf_counter()
        bc, ovf = ctx.bc_sources(src, mode)
        el = time.perf_counter() - t
        done = len(src)
        rate = done / el
        ctx.set_profiling(True)
        ctx.reset_timing()
        ctx.bc_sources(tree_batch(total // 2, total, size), mode)
        tm = ctx.timing()
        ctx.set_profiling(False)
        if tm.get("sweep_fwd_ms"):
            log(f"[bench]   sweep phases per tile: fwd {tm['sweep_fwd_ms']:.1f} bwd {tm['sweep_bwd_ms']:.1f} "
                f"asc {tm['sweep_asc_ms']:.1f} ms")
        log(f"[bench] {label}: {done} sources in {el:.3f}s -> {rate:.1f} sources/s; "
            f"phases {tm['scan_ms']:.1f} / {tm['dep_ms']:.1f} / {tm['reduce_ms']:.1f} ms")
        return rate, el, tm, ovf, done

    tuning = ctx.tuning()
    if not args.no_autotune:
        ta = time.time()
        tuning = ctx.autotune(tree_batch(0, total))
        log(f"[bench] autotune: {tuning} ({time.time() - ta:.1f}s)")
    scan_rate, scan_el, scan_tm, ovf1, _ = time_mode(bcx.MODE_BATCHED, "label-scan batched", B)
    Bs = min(args.sweep_batch or sweep_default, n)
    value, el, tm, ovf, done = time_mode(bcx.MODE_SWEEP, "bag-sweep batched", Bs)
    launches = -(-done // Bs)  # the steps, or fewer where they would pass a full BC
    ctx.set_tuning(batch=B)
    ovf = ovf or ovf1
    scan_ms = scan_tm["scan_ms"]
    dep_ms = scan_tm["dep_ms"]
    red_ms = scan_tm["reduce_ms"]
    S = 64
    # Roofline of the dominant kernel, the sweep tile kernel: algorithmic bytes
    # of one launch over its duration (HIP events on the context's stream).
    bags = int(st["bag_entries"])
    parent, _ = ix.tree()
    inner = int(np.unique(parent[parent >= 0]).size)  # nodes with descendants
    sweep_ms = tm["scan_ms"]
    algo = sweep_algorithmic_bytes(n, bags, inner, int(st["tree_height"]), Bs)
    achieved = algo / (sweep_ms * 1e-3) / 1e9
    traffic = committed_traffic(Bs, args.workload)

    # Plain form: one source per launch sequence.
    ps = tree_perm[: args.plain_sources]
    ctx.bc_sources(ps[:2], bcx.MODE_PLAIN)
    t = time.perf_counter()
    ctx.bc_sources(ps, bcx.MODE_PLAIN)
    plain = len(ps) / (time.perf_counter() - t)
    log(f"[bench] plain: {plain:.1f} sources/s ({value / plain:.2f}x batched/plain)")

    rq = None
    if not args.no_rq:
        rq = _rq_bench(bcx, gg, ctx, g, args.rq_requests, Bs)

    cpu = None
    if not args.no_cpu:
        cpu = _cpu_baseline(g, n, args.cpu_seconds)

    h2h = None
    if not args.no_h2h:
        ctx.close()  # its sweep buffers hold most of the device memory
        h2h = _h2h_bench(bcx, G, ix, tree_perm, n)

    out = {
        "metric": "sources/sec for full BC (batched vs plain) + RQ p50/p99 latency, 1 MI355X",
        "value": round(value, 2),
        "unit": "sources/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * el / launches, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32+u64 (labels, counts) / f64 (dependencies, BC)",
        "data": "synthetic (seeded road-like Delaunay graph, generated in-process)",
        "config": {
            "workload": f"{args.workload}: MHL full BC (bag-sweep form), n={n}, arcs={g.arcs}, "
                        f"{Bs} sources per batched launch",
            "n": n, "arcs": int(g.arcs), "sweep_batch": Bs, "labelscan_batch": B, "tile": S,
            "tchunk": tuning["tchunk"], "dep_threads": tuning["bp_threads"],
            "tree_height": st["tree_height"], "tree_width": st["tree_width"],
            "label_entries": st["label_entries"], "bag_entries": bags, "inner_nodes": inner,
            "Lbar": round(lbar, 2),
            "labelscan_sources_per_s": round(scan_rate, 2),
            "sweep_phase_ms": [round(tm["scan_ms"], 3), round(tm["dep_ms"], 3), round(tm["reduce_ms"], 3)],
            "plain_sources_per_s": round(plain, 2),
            "batched_over_plain": round(value / plain, 2),
            "full_bc_seconds_est": round(n / value, 2),
            "scan_ms_per_batch": round(scan_ms, 3), "dep_ms_per_batch": round(dep_ms, 3),
            "reduce_ms_per_batch": round(red_ms, 3),
            "count_overflow": bool(ovf),
            "h2h_scan": h2h,
        },
        "roofline": {
            "bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": None if traffic is None else round(traffic, 0),
            "kernel": SWEEP_KERNEL, "kernel_ms": round(sweep_ms, 3),
            "algorithmic_bytes": round(algo, 0),
            "traffic_GBs": None if traffic is None else round(traffic / (sweep_ms * 1e-3) / 1e9, 2),
        },
        "cpu_baseline": cpu,
        "rq": rq,
    }
    print(json.dumps(out), flush=True)
    return 0


def _h2h_bench(bcx, G, ix, tree_perm, n: int, sources: int = 512) -> dict:
    """QBC's distance-only path on the same index: H2H labels built on the
    host, then the device separator scan for `sources` sources (one block of
    the tree order), timed by HIP events around its kernels."""
    t0 = time.perf_counter()
    ix.build_h2h()
    build_s = time.perf_counter() - t0
    ctx = bcx.Context(G, ix, 0)
    src = tree_perm[(n - sources) // 2:(n - sources) // 2 + sources]
    ctx.h2h_distances(src[:8])  # warm-up
    ctx.set_profiling(True)
    ctx.reset_timing()
    t0 = time.perf_counter()
    ctx.h2h_distances(src)
    wall = time.perf_counter() - t0
    tm = ctx.timing()
    ctx.set_profiling(False)
    ctx.close()
    out = {"sources": int(len(src)), "label_build_s": round(build_s, 2),
           "kernel_ms": round(tm["scan_ms"], 3),
           "sources_per_s": round(len(src) / (tm["scan_ms"] * 1e-3), 2),
           "with_copy_back_sources_per_s": round(len(src) / wall, 2)}
    log(f"[bench] h2h distance scan: {out}")
    return out


def _rq_bench(bcx, gg, ctx, g, nreq: int, batch: int) -> dict:
    """Continuous-batching RQ serving on the bag-sweep form: Poisson arrivals,
    admission by budget; every request rides the shared cyclic sweep for one
    full cycle of batches."""
    n = g.n
    ctx.set_tuning(sweep_batch=batch)
    row = np.zeros(n + 1, np.int64)
    np.add.at(row, g.u.astype(np.int64) + 1, 1)
    row = np.cumsum(row)
    sets = gg.query_sets(n, nreq, seed=4, adjacency=(row, g.v))
    free, _ = ctx.mem_info()
    srv = bcx.RQServer(ctx, int(free * 0.25))
    # Saturating arrival rate: mean gap = one batch time / 2.
    rng_gap = gg.uniform01(4, 30, nreq)
    t_batch = None
    arrivals = []
    submitted = {}
    done = {}
    t0 = time.perf_counter()
    pending = list(range(nreq))
    next_arrival = 0.0
    while len(done) < nreq:
        now = time.perf_counter() - t0
        while pending and now >= next_arrival:
            i = pending[0]
            rid = srv.submit(sets[i])
            if rid is None:
                break
            pending.pop(0)
            submitted[rid] = (i, time.perf_counter())
            gap = -np.log(1.0 - rng_gap[i]) * (t_batch or 0.01) * 0.5
            next_arrival = now + gap
        ts = time.perf_counter()
        srv.step()
        t_step = time.perf_counter() - ts
        t_batch = t_step if t_batch is None else 0.9 * t_batch + 0.1 * t_step
        for rid in list(submitted):
            r = srv.poll(rid)
            if r is not None:
                i, ta = submitted.pop(rid)
                done[i] = time.perf_counter() - ta
        if not submitted and pending:
            next_arrival = min(next_arrival, time.perf_counter() - t0)
    wall = time.perf_counter() - t0
    lat = np.array([done[i] for i in range(nreq)])
    qs = np.array([len(s) for s in sets])
    res = {"requests": nreq, "p50_s": round(float(np.percentile(lat, 50)), 4),
           "p99_s": round(float(np.percentile(lat, 99)), 4),
           "p99_over_p50": round(float(np.percentile(lat, 99) / np.percentile(lat, 50)), 3),
           "requests_per_s": round(nreq / wall, 4), "queries_per_s": round(float(qs.sum()) / wall, 1),
           "q_sizes": qs.tolist(), "batch_seconds": round(t_batch or 0.0, 4),
           "form": "bag-sweep", "sources_per_batch": batch}
    log(f"[bench] rq: {res}")
    return res


def _cpu_baseline(g, n: int, budget_s: float) -> dict:
    """Oracle Dijkstra-Brandes (one thread) on a bounded sample of sources."""
    import oracle
    from gpu_betweennesscentrality_amd import graphgen as gg

    csr = oracle.CSR.from_arcgraph(g)
    src = gg.permutation(9, 1, n).astype(np.int32)
    t = time.perf_counter()
    oracle.brandes(csr, src[:2])
    per = max(1e-4, (time.perf_counter() - t) / 2)
    k = int(max(4, min(n, budget_s / per)))
    t = time.perf_counter()
    oracle.brandes(csr, src[:k])
    el = time.perf_counter() - t
    res = {"value": round(k / el, 3), "unit": "sources/s", "cores": 1, "kind": "port",
           "sample": f"{k} seeded random sources of the same graph, Dijkstra + Brandes "
                     f"(oracle/oracle.c), {el:.1f}s"}
    log(f"[bench] cpu baseline: {res}")
    return res


if __name__ == "__main__":
    sys.exit(main())
