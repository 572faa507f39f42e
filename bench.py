"""Benchmark of the betweenness hot path on one MI355X.

A step is one bag-sweep launch over one batch of source vertices (524,288 on
MI355X: 32 tiles of 64 sources per compute unit, drawing on two state slots
per unit; fewer where device memory caps the launch, or as the autotuner
picks) on a seeded road-like graph with its MHL index resident in HBM.  W untimed steps, then exactly K timed steps as one full-BC
call: K launches (tile kernel + reduction over tiles), then the shortcut
unpacking once; the sources walk the tree order cyclically.  `value` is
sources/s over that call.  Beside it: a full BC of every source, the
label-scan form (the reference's method) and the plain one-source-per-launch
form, the roofline of the tile kernel, self-checks of the timed result, RQ
serving at saturation and the CPU baseline (the reference's method on every
host core).

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
READ_KERNEL = "stream_read_kernel"
READ_BYTES = 4 << 30  # the --profile-only streaming read, larger than the 256 MB MALL
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
    (8 * (inner - height)).  Count slots (written and read only where a path
    count is too large for its distance word) and the quotients of further
    tight entries are data-dependent and not counted.
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
    out = (f"| counter | sweep kernel (KiB) | 1 GiB copy (KiB) | scale | bytes per launch |\n"
           f"|---|---|---|---|---|\n"
           f"| FETCH_SIZE | {fk:.0f} | {fc:.0f} | {ff:.3f} | {fetch:.4e} |\n"
           f"| WRITE_SIZE | {wk:.0f} | {wc:.0f} | {wf:.3f} | {write:.4e} |\n")
    try:
        fr = pick(f, READ_KERNEL)
        out += (f"\nCross-check of the FETCH_SIZE scale: the 4 GiB streaming read (`{READ_KERNEL}`) reads "
                f"{fr:.0f} KiB, scale {(READ_BYTES / 1024.0) / fr:.3f}.\n")
    except SystemExit:
        pass
    return out + f"\ntraffic_bytes_per_launch: {fetch + write:.6e}\n"


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


def max_rel(a, b) -> float:
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), 1.0))) if len(a) else 0.0


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=6, help="timed bag-sweep launches")
    ap.add_argument("--warmup", type=int, default=1, help="untimed bag-sweep launches before them")
    ap.add_argument("--workload", default="road320k")
    ap.add_argument("--batch", type=int, default=4096, help="label-scan sources per batch")
    ap.add_argument("--sweep-batch", type=int, default=0,
                    help="bag-sweep sources per launch (0: the library default, 32 tiles of 64 per CU)")
    ap.add_argument("--labelscan-full-max-n", type=int, default=400_000,
                    help="run the label scan over every source (the cross-check of the full BC) up to this n")
    ap.add_argument("--plain-sources", type=int, default=64)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-rq", action="store_true")
    ap.add_argument("--no-h2h", action="store_true", help="skip the QBC distance-only (H2H) scan")
    ap.add_argument("--rq-requests", type=int, default=200)
    ap.add_argument("--rq-slots", type=int, default=16,
                    help="RQ budget: the engine's fixed memory plus this many mean-size requests")
    ap.add_argument("--rq-load", type=float, default=1.25,
                    help="offered RQ load as a multiple of the engine's capacity")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--profile-only", action="store_true",
                    help="one label-scan batch, one sweep launch and a 1 GiB copy (for rocprofv3)")
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
    if args.steps < 1 or args.warmup < 0:
        print("--steps must be at least 1 and --warmup at least 0", file=sys.stderr)
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
        f"entries={st['label_entries']} bags={st['bag_entries']} Lbar={lbar:.1f} build={st['build_seconds']:.1f}s")
    t0 = time.time()
    ctx = bcx.Context(G, ix, 0)
    ctx_s = time.time() - t0
    n = g.n
    B = min(args.batch, n)
    ctx.set_tuning(batch=B)
    if args.sweep_batch:
        ctx.set_tuning(sweep_batch=args.sweep_batch)
    tree_perm = ix.tree_order()

    def cyc(start: int, count: int) -> np.ndarray:
        """`count` sources from position `start` of the tree order, cyclically."""
        return tree_perm[(start + np.arange(count, dtype=np.int64)) % n].astype(np.int32)

    W, K = args.warmup, args.steps
    checks = {}
    tuning = ctx.tuning()
    if not args.no_autotune and not args.profile_only:
        ta = time.time()
        tuning = ctx.autotune(cyc(0, B))
        if args.sweep_batch:  # an explicit launch size wins over the tuned one
            ctx.set_tuning(sweep_batch=args.sweep_batch)
        log(f"[bench] autotune: {tuning} ({time.time() - ta:.1f}s)")
    Bs = ctx.sweep_capacity()  # sources per launch after the memory cap
    if args.profile_only:
        # The same bag-sweep launches as the timed call (W + K, one job),
        # one label-scan batch and a 1 GiB copy, for a kernel trace.
        ctx.bc_sources(cyc(n // 3, B), bcx.MODE_BATCHED)
        ctx.bc_sources(cyc(0, (W + K) * Bs), bcx.MODE_SWEEP)
        ms = ctx.copy_probe(COPY_BYTES)
        rms = ctx.read_probe(READ_BYTES)
        log(f"[bench] profile-only done ({W + K} launches of {Bs} sources); 1 GiB device copy {ms:.3f} ms; "
            f"4 GiB streaming read {rms:.3f} ms ({READ_BYTES / rms / 1e6:.0f} GB/s)")
        return 0

    # ---- label scan (the reference's method): rate, and the cross-check ------
    ls_full = n <= args.labelscan_full_max_n
    ls_src = tree_perm.astype(np.int32) if ls_full else cyc(n // 5, 4 * B)
    ctx.bc_sources(cyc(0, B), bcx.MODE_BATCHED)  # warm-up
    t = time.perf_counter()
    ls_bc, ls_ovf = ctx.bc_sources(ls_src, bcx.MODE_BATCHED)
    ls_el = time.perf_counter() - t
    ls_rate = len(ls_src) / ls_el
    log(f"[bench] label scan: {len(ls_src)} sources in {ls_el:.2f}s -> {ls_rate:.1f} sources/s")

    # ---- headline: W untimed, then K timed bag-sweep launches ---------------
    if W:
        ctx.bc_sources(cyc(0, W * Bs), bcx.MODE_SWEEP)
    src = cyc(W * Bs, K * Bs)
    ctx.reset_timing()
    t = time.perf_counter()
    timed_bc, ovf = ctx.bc_sources(src, bcx.MODE_SWEEP)
    el = time.perf_counter() - t
    launches = ctx.timing()["scan_launches"]
    value = K * Bs / el
    log(f"[bench] bag sweep: {K} launches of {Bs} sources in {el:.3f}s -> {value:.1f} sources/s "
        f"({1e3 * el / K:.2f} ms per launch)")
    checks["timed_call_launches"] = {"want": K, "got": int(launches), "ok": int(launches) == K}

    # Tile-kernel time per launch: the same K launches again, each bracketed
    # by HIP events on the context's stream (the time depends on where in
    # the tree order a launch's sources lie, so the mean is over these).
    ctx.set_profiling(True)
    ctx.reset_timing()
    prof_bc, _ = ctx.bc_sources(src, bcx.MODE_SWEEP)
    tm = ctx.timing()
    ctx.set_profiling(False)
    pl = max(1, int(tm["scan_launches"]))
    kernel_ms = tm["scan_ms"] / pl
    log(f"[bench] sweep tile kernel {kernel_ms:.2f} ms per launch of {Bs} (mean of {pl}); per tile fwd "
        f"{tm['sweep_fwd_ms'] / pl:.1f} bwd {tm['sweep_bwd_ms'] / pl:.1f} asc {tm['sweep_asc_ms'] / pl:.1f} ms; "
        f"reduce {tm['reduce_ms'] / pl:.2f} ms")
    checks["timed_equals_eager_bitwise"] = {"ok": bool(np.array_equal(timed_bc.view(np.uint64),
                                                                      prof_bc.view(np.uint64)))}

    # ---- a full BC of every source -----------------------------------------
    t = time.perf_counter()
    full_bc, full_ovf = ctx.full_bc(bcx.MODE_SWEEP)
    full_first = time.perf_counter() - t
    t = time.perf_counter()
    full_bc2, _ = ctx.full_bc(bcx.MODE_SWEEP)
    full_el = time.perf_counter() - t
    log(f"[bench] full BC (bag sweep): {full_el:.3f}s ({n / full_el:.1f} sources/s; first call {full_first:.3f}s)")
    checks["full_bc_repeat_bitwise"] = {"ok": bool(np.array_equal(full_bc.view(np.uint64),
                                                                  full_bc2.view(np.uint64)))}

    # ---- self-checks of the timed result -------------------------------------
    # The timed sources are q whole passes over the tree order plus r more:
    # BC(timed) = q * BC(all) + BC(those r), within rounding.
    q, r = divmod(K * Bs, n)
    rest = ctx.bc_sources(cyc(W * Bs + q * n, r), bcx.MODE_SWEEP)[0] if r else np.zeros(n)
    e = max_rel(timed_bc, q * full_bc + rest)
    checks["timed_equals_passes_plus_rest"] = {"max_rel": e, "tol": 1e-9, "ok": e <= 1e-9}
    # The sweep against the reference's method (label scan) on the same sources.
    sw_ls = full_bc if ls_full else ctx.bc_sources(ls_src, bcx.MODE_SWEEP)[0]
    e = max_rel(sw_ls, ls_bc)
    checks["sweep_equals_label_scan"] = {"sources": int(len(ls_src)), "max_rel": e, "tol": 1e-9, "ok": e <= 1e-9}
    checks["no_count_overflow"] = {"ok": not (ovf or full_ovf or ls_ovf)}
    for k, v in checks.items():
        log(f"[bench] check {k}: {v}")

    # ---- roofline ------------------------------------------------------------
    bags = int(st["bag_entries"])
    parent, _ = ix.tree()
    inner = int(np.unique(parent[parent >= 0]).size)  # nodes with descendants
    algo = sweep_algorithmic_bytes(n, bags, inner, int(st["tree_height"]), Bs)
    achieved = algo / (kernel_ms * 1e-3) / 1e9
    traffic = committed_traffic(Bs, args.workload)
    # Label-scan form against its own byte model, B_src(64) with the measured
    # mean DAG arcs per source.
    d_rows = ctx.query_rows(cyc(n // 7, 4), bcx.MODE_BATCHED)[0].astype(np.int64)
    dbar = float(np.mean([np.count_nonzero(d[g.v] == d[g.u] + g.w) for d in d_rows]))
    bsrc = 16.0 * n * lbar / 64 + 48.0 * n + 4.0 * g.arcs + 8.0 * dbar

    # ---- plain form: one source per launch sequence ---------------------------
    ps = cyc(n // 3, args.plain_sources)
    ctx.bc_sources(ps[:2], bcx.MODE_PLAIN)
    t = time.perf_counter()
    ctx.bc_sources(ps, bcx.MODE_PLAIN)
    plain = len(ps) / (time.perf_counter() - t)
    log(f"[bench] plain: {plain:.1f} sources/s")

    rq = None
    if not args.no_rq:
        rq = _rq_bench(bcx, gg, ctx, g, args.rq_requests, args.rq_slots, args.rq_load)

    cpu = None
    if not args.no_cpu:
        cpu = _cpu_baseline(g, ix, n, args.cpu_seconds, st)

    h2h = None
    if not args.no_h2h:
        ctx.close()  # its sweep buffers hold most of the device memory
        h2h = _h2h_bench(bcx, G, ix, tree_perm, n)

    ok = all(v["ok"] for v in checks.values())
    out = {
        "metric": "sources/sec for full BC (batched vs plain) + RQ p50/p99 latency, 1 MI355X",
        "value": round(value, 2),
        "unit": "sources/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": round(1e3 * el / K, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32+u64 (labels, counts) / f64 (dependencies, BC)",
        "data": "synthetic (seeded road-like Delaunay graph, generated in-process)",
        "config": {
            "workload": f"{args.workload}: MHL full BC, bag-sweep form, n={n}, arcs={g.arcs}; a step is one launch "
                        f"of {Bs} sources, the timed call {K} launches plus one unpack",
            "n": n, "arcs": int(g.arcs), "sources_per_launch": Bs, "timed_launches": int(launches),
            "timed_sources": K * Bs, "timed_seconds": round(el, 4),
            "full_bc_sources_per_s": round(n / full_el, 2), "full_bc_seconds": round(full_el, 4),
            "full_bc_first_call_seconds": round(full_first, 4),
            "index_build_seconds": round(st["build_seconds"], 3), "context_create_seconds": round(ctx_s, 3),
            "labelscan_batch": B, "tchunk": tuning["tchunk"], "dep_threads": tuning["bp_threads"],
            "tree_height": st["tree_height"], "tree_width": st["tree_width"],
            "label_entries": st["label_entries"], "bag_entries": bags, "inner_nodes": inner,
            "Lbar": round(lbar, 2), "Dbar": round(dbar, 1),
            "labelscan_sources_per_s": round(ls_rate, 2), "labelscan_sources": int(len(ls_src)),
            "plain_sources_per_s": round(plain, 2),
            "sweep_over_plain": round(value / plain, 2),
            "labelscan_over_plain": round(ls_rate / plain, 2),
            "sweep_over_labelscan": round(value / ls_rate, 2),
            "count_overflow": bool(ovf),
            "h2h_scan": h2h,
        },
        "checks": checks,
        "roofline": {
            "bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": None if traffic is None else round(traffic, 0),
            "model": "bag-sweep lower bound per source: 8*bag_entries + 4*n + 12*inner_nodes "
                     "+ 8*(inner_nodes - tree_height) bytes",
            "kernel": SWEEP_KERNEL, "kernel_ms": round(kernel_ms, 3), "sources_per_launch": Bs,
            "algorithmic_bytes": round(algo, 0),
            "traffic_GBs": None if traffic is None else round(traffic / (kernel_ms * 1e-3) / 1e9, 2),
            "labelscan": {
                "model": "B_src(S) = 16*n*Lbar/S + 48*n + 4*arcs + 8*Dbar, S = 64",
                "bytes_per_source": round(bsrc, 0),
                "achieved_GBs": round(ls_rate * bsrc / 1e9, 2),
                "frac": round(ls_rate * bsrc / 1e9 / HBM_PEAK_GBS, 4)},
        },
        "cpu_baseline": cpu,
        "rq": rq,
    }
    print(json.dumps(out), flush=True)
    if not ok:
        log("[bench] a self-check failed: " + ", ".join(k for k, v in checks.items() if not v["ok"]))
        return 1
    return 0


def _h2h_bench(bcx, G, ix, tree_perm, n: int, sources: int = 512) -> dict:
    """QBC's distance-only path on the same index: H2H labels, optimizeVS
    separators and P2H branch tables built on the host, then the device
    separator scan for `sources` sources (one block of the tree order), timed
    by HIP events around its kernels; the mean scan length with and without
    the P2H tables beside it."""
    t0 = time.perf_counter()
    ix.build_h2h()
    build_s = time.perf_counter() - t0
    scan_sep, scan_p2h = ix.h2h_scan_stats(1 << 18, 1)
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
           "mean_scan_optimizeVS": round(scan_sep, 2), "mean_scan_p2h": round(scan_p2h, 2),
           "kernel_ms": round(tm["scan_ms"], 3),
           "sources_per_s": round(len(src) / (tm["scan_ms"] * 1e-3), 2),
           "with_copy_back_sources_per_s": round(len(src) / wall, 2)}
    log(f"[bench] h2h distance scan: {out}")
    return out


def _rq_bench(bcx, gg, ctx, g, nreq: int, slots: int, load: float) -> dict:
    """Continuous-batching RQ serving at saturation on the bag-sweep form.

    Requests (Q log-uniform in [1e3, 5e4], uniform or BFS-ball sets) arrive
    as a seeded Poisson stream at `load` times the engine's capacity.  The
    engine's budget is its fixed memory plus `slots` mean-size requests, so
    admission binds: a request that does not fit gets BCX_EBUSY and waits in
    arrival order.  Every admitted request rides the shared cyclic sweep for
    one full cycle.  Latency = completion - arrival, split into queue wait
    (arrival to admission) and service (admission to completion)."""
    n = g.n
    row = np.zeros(n + 1, np.int64)
    np.add.at(row, g.u.astype(np.int64) + 1, 1)
    row = np.cumsum(row)
    sets = gg.query_sets(n, nreq, seed=4, adjacency=(row, g.v))
    probe = bcx.RQServer(ctx, 1 << 40)
    fixed, _ = probe.memory()
    need = np.array([probe.bytes_needed(len(s)) for s in sets], np.float64)
    probe.close()
    budget = int(fixed + max(slots * need.mean(), need.max()))  # the largest request always fits alone
    srv = bcx.RQServer(ctx, budget)
    # One step is one sweep launch; a request is served after nb steps.
    ts = time.perf_counter()
    rid = srv.submit(sets[0][:1])
    steps_cycle = 0
    while srv.poll(rid) is None:
        srv.step()
        steps_cycle += 1
    cycle_s = time.perf_counter() - ts
    capacity = slots / cycle_s  # requests per second with `slots` in flight
    rate = load * capacity
    gaps = -np.log(1.0 - gg.uniform01(4, 30, nreq)) / rate
    arrive = np.cumsum(gaps)
    admit = np.full(nreq, np.nan)
    done = np.full(nreq, np.nan)
    inflight = {}
    nxt = 0
    ebusy = 0
    steps = 0
    max_inflight = 0
    t0 = time.perf_counter()
    while np.isnan(done).any():
        now = time.perf_counter() - t0
        while nxt < nreq and arrive[nxt] <= now:
            r = srv.submit(sets[nxt])
            if r is None:
                ebusy += 1
                break
            admit[nxt] = time.perf_counter() - t0
            inflight[r] = nxt
            nxt += 1
        max_inflight = max(max_inflight, len(inflight))
        if not inflight:
            time.sleep(max(0.0, arrive[nxt] - (time.perf_counter() - t0)))
            continue
        srv.step()
        steps += 1
        for r in list(inflight):
            if srv.poll(r) is not None:
                done[inflight.pop(r)] = time.perf_counter() - t0
    wall = time.perf_counter() - t0
    srv.close()
    lat, wait, serv = done - arrive, admit - arrive, done - admit
    qs = np.array([len(s) for s in sets])

    def pct(x, p):
        return round(float(np.percentile(x, p)), 4)

    res = {"requests": nreq, "form": "bag-sweep", "steps": steps, "cycles": round(steps / steps_cycle, 2),
           "steps_per_cycle": steps_cycle, "cycle_s": round(cycle_s, 4),
           "budget_bytes": budget, "fixed_bytes": int(fixed), "slots": slots, "max_in_flight": max_inflight,
           "ebusy": ebusy, "offered_per_s": round(rate, 3), "capacity_per_s": round(capacity, 3),
           "p50_s": pct(lat, 50), "p99_s": pct(lat, 99), "p99_over_p50": round(pct(lat, 99) / pct(lat, 50), 3),
           "wait_p50_s": pct(wait, 50), "wait_p99_s": pct(wait, 99),
           "service_p50_s": pct(serv, 50), "service_p99_s": pct(serv, 99),
           "service_p99_over_p50": round(pct(serv, 99) / pct(serv, 50), 3),
           "requests_per_s": round(nreq / wall, 3), "queries_per_s": round(float(qs.sum()) / wall, 1),
           "q_min": int(qs.min()), "q_median": int(np.median(qs)), "q_max": int(qs.max())}
    log(f"[bench] rq: {res}")
    return res


def _cpu_baseline(g, ix, n: int, budget_s: float, st: dict) -> dict:
    """The reference's method on the host cores: the index's label rows
    (built by the host index builder, timed separately), then per source the
    label query of every target and the layered Brandes pass
    (oracle.label_bc, OpenMP over sources), on a bounded sample of the same
    workload; Dijkstra-Brandes (oracle.brandes) timed beside it."""
    import os

    import oracle
    from gpu_betweennesscentrality_amd import graphgen as gg

    threads = oracle.default_threads()
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    csr = oracle.CSR.from_arcgraph(g)
    tree = ix.export()
    src = gg.permutation(9, 1, n).astype(np.int32)

    def rate(fn, seconds):
        t = time.perf_counter()
        fn(src[:threads])
        per = max(1e-4, (time.perf_counter() - t) / threads)
        k = int(max(threads, min(n, seconds / per)))
        k -= k % threads
        t = time.perf_counter()
        fn(src[:k])
        el = time.perf_counter() - t
        return k / el, k, el

    lr, lk, lel = rate(lambda s: oracle.label_bc(csr, tree, s, threads), budget_s)
    dr, dk, dlel = rate(lambda s: oracle.brandes(csr, s, threads), budget_s / 3)
    res = {"value": round(lr, 3), "unit": "sources/s", "cores": threads, "kind": "port",
           "sample": f"{lk} seeded random sources of the same graph: label queries over the host index's rows "
                     f"to every target, then layered Brandes (oracle/oracle.c orc_label_bc), {lel:.1f}s "
                     f"on {threads} threads",
           "cpu_model": model, "nproc": os.cpu_count(), "threads": threads,
           "index_build_s": round(float(st["build_seconds"]), 3),
           "full_bc_seconds_est": round(n / lr, 1),
           "dijkstra_brandes_sources_per_s": round(dr, 3),
           "dijkstra_brandes_sample": f"{dk} sources, {dlel:.1f}s on {threads} threads"}
    log(f"[bench] cpu baseline: {res}")
    return res


if __name__ == "__main__":
    sys.exit(main())
