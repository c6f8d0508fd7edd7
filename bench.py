"""Benchmark: exact 4-clique count on a seeded RMAT graph, one MI355X.

Workload (default): RMAT scale 24, edge factor 16, a/b/c/d = .57/.19/.19/.05,
seed 1; symmetrised, deduplicated, self-loops and isolated vertices dropped,
degree-oriented; uint32 CSR written to and loaded back from a .dgraph file.
The graph is resident in HBM before timing starts.

A step is one full exact count of the graph's 4-cliques.  The persistent
form (one hipGraph launch of the degree-bucketed kernels) is the headline;
the plain level-by-level form is timed beside it.  `value` is the persistent
form's seconds per count; the count is checked against the committed
oracle count of the same graph (tests/golden/config_counts.json, matched by
CSR digest).  The roofline uses the bytes the kernels must stream,
4 (m + sum over edges (u, v) of d+(v)), over the HIP-event time of the
launch.  The CPU baseline is the oracle's restatement of the kautomine
four_clique schedule on a bounded sample of roots, on this host's cores,
with its time projected to the whole count by merge-model bytes.

    python bench.py [--steps K] [--warmup W] [--scale 24] [--plain-steps 1]

Other workloads (each prints its own JSON line; the default is the one above):
    --workload census   induced 3/4-vertex census of a seeded Chung-Lu graph
                        (n 4.85M, 43M draws, exponent 2.3, degree cap 20000)
    --workload batch    many seeded ER/RMAT graphs (seed 1000+i, n ~ U[2^10,
                        2^18], average degree ~ U[4, 64]), triangle, 4- and
                        5-clique of every graph in one hipGraph launch
    --workload gen      RMAT generation + clean-up + orientation on the host
                        and on the GPU (kz_gpu_gen_rmat), same digest
    --workload fsm      labelled frequent subgraph mining (kautomine three_fsm)
                        on the census graph with uniform labels (--labels,
                        --threshold)
    --workload apps     the unchanged three_motif apps (kautomine, kgraphpi) on
                        the census graph through the application boundary
    --workload serve    a seeded Poisson stream of queries through kz_server:
                        triangle and 4-clique on a large RMAT graph, 5-clique
                        and motif census on a pool of batch-workload graphs,
                        at several arrival rates; p50/p99 latency
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import tempfile
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
METRIC = "exact 4-clique count: s & GB/s, batched/persistent vs plain, 1 MI355X"


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_baseline(g, budget_s: float, threads: int) -> dict:
    """Oracle kautomine four_clique on a strided root sample sized to take
    about `budget_s` seconds; bytes of the same sample from the oracle."""
    import numpy as np

    import oracle

    n = g.num_vertices
    rp, col = g.rowptr, g.col
    frac = 1e-3
    while True:
        stride = max(1, int(round(1.0 / frac)))
        roots = np.arange(0, n, stride, dtype=np.uint32)
        t0 = time.perf_counter()
        oracle.run_app("kautomine_four_clique", rp, col, roots=roots, threads=threads)
        dt = time.perf_counter() - t0
        if dt >= 0.25 * budget_s or stride == 1:
            break
        frac = min(1.0, frac * max(2.0, 0.5 * budget_s / max(dt, 1e-3)))
    b = oracle.clique_bytes(rp, col, kmax=4, threads=threads, roots=roots)
    t0 = time.perf_counter()
    _, cnt = oracle.run_app("kautomine_four_clique", rp, col, roots=roots, threads=threads)
    dt = time.perf_counter() - t0
    return {
        "value": None,  # set by the caller: the sample's time projected to the whole count
        "unit": "s",
        "cores": threads,
        "kind": "port",
        "sample": f"every {stride}-th root vertex ({len(roots)} of {n}) of the same graph, "
                  f"{cnt[0]} 4-cliques, {b['B4']} merge-model bytes, {dt:.2f} s",
        "sample_seconds": dt,
        "sample_bytes": b["B4"],
        "sample_merge_model_GBps": b["B4"] / dt / 1e9,
    }


def pool_graph(kz, i: int, gpu: bool = False, orient: bool = False):
    """Graph i of the batch workload (openkhuzdulgraph_amd.workloads)."""
    from openkhuzdulgraph_amd import workloads as wl

    return wl.pool_graph(kz, i, gpu=gpu, orient=orient)


def base_line(metric: str, value, unit: str, steps: int, warmup: int, ms: float, higher: bool, config: dict) -> dict:
    return {"metric": metric, "value": value, "unit": unit, "n_gpus": 1, "steps": steps, "warmup": warmup,
            "ms_per_step": ms, "higher_is_better": higher, "scaling": "weak", "vs_baseline": None,
            "dtype": "uint32", "data": "synthetic", "config": config}


def bench_census(kz, args) -> int:
    t0 = time.perf_counter()
    from openkhuzdulgraph_amd import workloads as wl

    g = wl.chunglu_census(kz)
    log(f"generated {g} in {time.perf_counter() - t0:.1f} s")
    kz.init(0)
    dg = g.to_device()
    t0 = time.perf_counter()
    first = dg.motif_census()
    log(f"first call (GPU relabelling + census) {time.perf_counter() - t0:.2f} s")
    for _ in range(args.warmup):
        dg.motif_census()
    ms = []
    for _ in range(args.steps):
        c = dg.motif_census()
        if c != first:
            raise SystemExit("census changed between steps")
        ms.append(dg.last_stats["device_ms"])
    avg = sum(ms) / len(ms)
    check = None
    if args.check:
        import oracle

        threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or (os.cpu_count() or 1)
        t0 = time.perf_counter()
        want = oracle.motif_census(g.rowptr, g.col, threads=threads)
        check = {"equal": want == first, "oracle_seconds": time.perf_counter() - t0, "threads": threads}
        log(f"oracle check: {check}")
        if want != first:
            raise SystemExit("census differs from the oracle")
    from openkhuzdulgraph_amd._ffi import CENSUS_PATTERNS
    line = base_line("induced 3/4-vertex motif census: s, 1 MI355X", avg / 1e3, "s", args.steps, args.warmup, avg,
                     False, {"workload": "chunglu_4.85M_43Mdraws_census", "vertices": g.num_vertices,
                             "edges": g.num_entries // 2, "max_degree": g.max_degree, "seed": 1})
    line["counts"] = dict(zip(CENSUS_PATTERNS, first))
    line["oracle_check"] = check
    print(json.dumps(line), flush=True)
    return 0


def bench_fsm(kz, args) -> int:
    """Labelled frequent subgraph mining (kz_fsm_mine, kautomine three_fsm):
    the config-3-sized Chung-Lu graph built on the GPU with uniform labels,
    one full mining run per step; the CPU baseline is the oracle's literal
    restatement of the app on a small graph of the same family, with the
    GPU timed on that graph too."""
    import numpy as np
    kz.init(0)
    t0 = time.perf_counter()
    dg = kz.gpu_gen_chung_lu(4850000, 43000000, 2.3, 20000, 1)
    lab = np.random.default_rng(1).integers(0, args.labels, dg.num_vertices).astype(np.uint32)
    dg.set_labels(lab)
    nv, ne = dg.num_vertices, dg.num_entries
    log(f"built on the GPU + labelled in {time.perf_counter() - t0:.1f} s: {dg.num_vertices} vertices, "
        f"{dg.num_entries} entries, {args.labels} labels")
    first = dg.fsm(args.threshold, cap=1)
    for _ in range(args.warmup):
        dg.fsm(args.threshold, cap=1)
    ms, dev = [], []
    for _ in range(args.steps):
        r = dg.fsm(args.threshold, cap=1)
        if r["counts"] != first["counts"] or r["sums"] != first["sums"]:
            raise SystemExit("mining result changed between steps")
        ms.append(dg.last_stats["total_ms"])
        dev.append(dg.last_stats["device_ms"])
    avg = sum(ms) / len(ms)
    log(f"fsm: {avg:.1f} ms per run (device {sum(dev) / len(dev):.1f} ms), "
        f"{first['num_candidates']} candidates, frequent per level {first['counts']}")
    dg.release()
    cpu = None
    if not args.no_cpu:
        import oracle
        small = kz.gen_chung_lu(8000, 40000, 2.3, 800, 9)
        sl = np.random.default_rng(1).integers(0, args.labels, small.num_vertices).astype(np.uint32)
        st = args.threshold * small.num_vertices // nv or 1
        t0 = time.perf_counter()
        want = oracle.fsm(small.rowptr, small.col, sl, st, cap=1)
        cpu_s = time.perf_counter() - t0
        sd = small.to_device()
        sd.set_labels(sl)
        got = sd.fsm(st, cap=1)
        got = sd.fsm(st, cap=1)
        if got["counts"] != want["counts"] or got["sums"] != want["sums"]:
            raise SystemExit("GPU mining differs from the oracle on the CPU-baseline graph")
        cpu = {"value": cpu_s, "unit": "s", "cores": 1, "kind": "port",
               "sample": f"whole mining run on Chung-Lu n=8000, 40000 draws, {args.labels} labels, threshold {st} "
                         f"(the oracle enumerates embeddings literally, as the app does); the GPU takes "
                         f"{sd.last_stats['total_ms']:.1f} ms on it",
               "gpu_ms_same_graph": sd.last_stats["total_ms"]}
        sd.release()
        log(f"cpu baseline: {cpu}")
    line = base_line("labelled frequent subgraph mining (three_fsm, MNI support): s, 1 MI355X", avg / 1e3, "s",
                     args.steps, args.warmup, avg, False,
                     {"workload": "chunglu_4.85M_43Mdraws_fsm", "vertices": nv, "entries": ne,
                      "labels": args.labels, "threshold": args.threshold, "seed": 1})
    line["frequent"] = dict(zip(("vertex", "edge", "wedge", "triangle", "four_chain", "three_star"),
                                first["counts"]))
    line["support_sum"] = dict(zip(("vertex", "edge", "wedge", "triangle", "four_chain", "three_star"),
                                   first["sums"]))
    line["candidates"] = first["num_candidates"]
    line["cpu_baseline"] = cpu
    print(json.dumps(line), flush=True)
    return 0


def bench_batch(kz, args) -> int:
    t0 = time.perf_counter()
    kz.init(0)
    if args.host_gen:
        dgs = [pool_graph(kz, i, orient=True).to_device() for i in range(args.graphs)]
    else:
        dgs = [pool_graph(kz, i, gpu=True, orient=True) for i in range(args.graphs)]
    ent = sum(dg.num_entries for dg in dgs)
    log(f"generated {args.graphs} graphs ({ent} oriented entries) on the {'host' if args.host_gen else 'GPU'} "
        f"in {time.perf_counter() - t0:.1f} s")
    ks = [3, 4, 5]
    b = kz.Batch(dgs, ks)
    t0 = time.perf_counter()
    first = b.run()
    prep = time.perf_counter() - t0
    log(f"first launch (task lists, capture) {prep:.2f} s")
    for _ in range(args.warmup):
        b.run()
    ms = []
    for _ in range(args.steps):
        if not (b.run() == first).all():
            raise SystemExit("batch counts changed between launches")
        ms.append(b.last_stats["device_ms"])
    avg = sum(ms) / len(ms)
    check = None
    if args.check:
        import oracle

        idx = list(range(0, args.graphs, max(1, args.graphs // 16)))
        for i in idx:
            for j, k in enumerate(ks):
                if dgs[i].count_cliques(k, "persistent") != int(first[i, j]):
                    raise SystemExit(f"batch differs from the single-graph run: graph {i} k {k}")
        small = sorted(range(args.graphs), key=lambda i: dgs[i].num_entries)[:4]
        apps = {3: "kautomine_triangle", 4: "kautomine_four_clique", 5: "kautomine_five_clique"}
        for i in small:
            h = dgs[i].download()
            for j, k in enumerate(ks):
                if oracle.run_app(apps[k], h.rowptr, h.col)[1][0] != int(first[i, j]):
                    raise SystemExit(f"batch differs from the oracle: graph {i} k {k}")
        check = {"single_runs": len(idx), "oracle_graphs": len(small)}
        log(f"checks: {check}")
    q = args.graphs * len(ks)
    line = base_line("batched exact clique counts: queries/s, 1 MI355X", q / (avg / 1e3), "queries/s", args.steps,
                     args.warmup, avg, True, {"workload": f"batch_{args.graphs}_graphs_k345", "graphs": args.graphs,
                                              "oriented_entries": ent, "queries": q})
    line["prepare_seconds"] = prep
    line["totals"] = {f"k{k}": int(first[:, j].sum()) for j, k in enumerate(ks)}
    line["check"] = check
    print(json.dumps(line), flush=True)
    return 0


def bench_serve(kz, args) -> int:
    import numpy as np

    t0 = time.perf_counter()
    kz.init(0)
    if args.host_gen:
        dbig = kz.gen_rmat(args.scale, args.edge_factor, 0.57, 0.19, 0.19, args.seed).oriented().to_device()
        dsym = [pool_graph(kz, i).to_device() for i in range(args.graphs)]
        dori = [pool_graph(kz, i, orient=True).to_device() for i in range(args.graphs)]
    else:
        dbig = kz.gpu_gen_rmat(args.scale, args.edge_factor, 0.57, 0.19, 0.19, args.seed, orient=True)
        dsym = [pool_graph(kz, i, gpu=True) for i in range(args.graphs)]
        dori = [pool_graph(kz, i, gpu=True, orient=True) for i in range(args.graphs)]
    log(f"generated the large graph ({dbig.num_vertices} vertices, {dbig.num_entries} oriented entries) and "
        f"{args.graphs} pool graphs in {time.perf_counter() - t0:.1f} s")
    # query kinds: (graph, kind, k); large graph: triangle, 4-clique; pool:
    # 5-clique (oriented copy) and the motif census (symmetric copy)
    def kinds(i):
        return [(dbig, "clique", 3, "big"), (dbig, "clique", 4, "big"), (dori[i], "clique", 5, f"pool{i}"),
                (dsym[i], "census", 0, f"pool{i}")]

    rng = np.random.default_rng(2)
    budget = int(args.budget_gb * (1 << 30))
    # each query kind alone first: the reference result for the mix
    alone = {}
    with kz.Server(hbm_budget=budget, max_inflight=1) as srv:
        for i in range(args.graphs):
            for g, kind, k, name in kinds(i):
                key = (name, kind, k)
                if key in alone:
                    continue
                r = srv.poll(srv.submit(g, kind, k or 4), wait=True)
                if r["status"] != 0:
                    raise SystemExit(f"query {key} failed: {r}")
                alone[key] = r["counts"]
    log(f"alone results for {len(alone)} query kinds")
    rates = [float(x) for x in args.rates.split(",")]
    per_rate = []
    with kz.Server(hbm_budget=budget, max_inflight=8) as srv:
        for rate in rates:
            t_start = time.perf_counter()
            next_t = t_start
            outstanding, lat, done, mismatches = {}, [], 0, 0
            next_log = t_start + 30.0
            while True:
                now = time.perf_counter()
                if now >= next_log:
                    log(f"rate {rate}/s: {done} done, {len(outstanding)} outstanding after {now - t_start:.0f} s")
                    next_log += 30.0
                if now >= next_t and now - t_start < args.seconds:
                    i = int(rng.integers(0, args.graphs))
                    big_q = rng.random() < args.large_share
                    g, kind, k, name = kinds(i)[int(rng.integers(0, 2)) + (0 if big_q else 2)]
                    outstanding[srv.submit(g, kind, k or 4)] = (name, kind, k)
                    next_t += float(rng.exponential(1.0 / rate))
                    continue
                for t in list(outstanding):
                    r = srv.poll(t)
                    if r is None:
                        continue
                    key = outstanding.pop(t)
                    if r["status"] != 0 or r["counts"] != alone[key]:
                        mismatches += 1
                        if mismatches <= 10:
                            log(f"mismatch {key}: status {r['status']} counts {r['counts'][:3]} "
                                f"alone {alone[key][:3]} total_ms {r['total_ms']:.1f}")
                    lat.append(r["total_ms"])
                    done += 1
                if now - t_start >= args.seconds and not outstanding:
                    break
                time.sleep(0.0002)
            wall = time.perf_counter() - t_start
            lat.sort()
            entry = {"rate": rate, "queries": done, "queries_per_s": done / wall,
                     "p50_ms": lat[len(lat) // 2] if lat else None,
                     "p99_ms": lat[min(len(lat) - 1, int(0.99 * len(lat)))] if lat else None,
                     "mismatches": mismatches}
            per_rate.append(entry)
            log(f"rate {rate}/s: {entry}")
            if mismatches:
                raise SystemExit(f"{mismatches} queries differ from their alone results")
    best = max(per_rate, key=lambda e: e["queries_per_s"])
    line = base_line("serving exact counts: queries/s and p50/p99 latency, 1 MI355X", best["queries_per_s"],
                     "queries/s", len(rates), 0, best["p50_ms"], True,
                     {"workload": f"serve_rmat{args.scale}_ef{args.edge_factor}_pool{args.graphs}",
                      "large_graph_vertices": dbig.num_vertices, "large_graph_oriented_edges": dbig.num_entries,
                      "pool_graphs": args.graphs, "budget_bytes": budget, "max_inflight": 8,
                      "mix": f"{args.large_share:.2f} large-graph queries (triangle, 4-clique), the rest pool "
                             "queries (5-clique, motif census), each pair uniform"})
    line["per_rate"] = per_rate
    print(json.dumps(line), flush=True)
    return 0


def bench_apps(kz, args) -> int:
    """The unchanged kautomine and kgraphpi three_motif apps (compiled from
    the reference sources against the Khuzdul API headers) on the census
    graph's .dgraph, through the application boundary: probe, plain-form
    GPU count and the app's own timing; their counts are checked against
    the motif census of the same graph (kautomine prints the induced wedges
    and the triangles, kgraphpi the triangles and the non-induced wedges)."""
    import re
    import subprocess

    import numpy as np

    from openkhuzdulgraph_amd import workloads as wl

    apps_dir = ROOT / "openkhuzdulgraph_amd" / "khuzdul" / "_apps"
    if not apps_dir.is_dir():
        raise SystemExit(f"{apps_dir} missing: the apps are built where the reference sources are")
    t0 = time.perf_counter()
    g = wl.chunglu_census(kz)
    log(f"generated {g} in {time.perf_counter() - t0:.1f} s")
    kz.init(0)
    dg = g.to_device()
    census = dg.motif_census()
    census_ms = dg.last_stats["device_ms"]
    dg.release()
    wedges, tri = census[0], census[1]
    deg = g.degrees().astype(np.int64)
    noninduced_wedges = int((deg * (deg - 1) // 2).sum())
    want = {"kautomine_three_motif": [tri, wedges], "kgraphpi_three_motif": [tri, noninduced_wedges]}
    results = {}
    with tempfile.TemporaryDirectory() as td:
        path = os.path.join(td, "chunglu.dgraph")
        g.store(path)
        for app, expect in want.items():
            out = subprocess.run([str(apps_dir / app), path, "1", str(args.steps), "true"], capture_output=True,
                                 text=True, timeout=1200)
            if out.returncode != 0:
                raise SystemExit(f"{app} failed: {out.stderr[-2000:]}")
            counts = {}
            for m in re.finditer(r"Pattern count (\d+): (\d+)", out.stdout):
                counts[int(m.group(1))] = int(m.group(2))
            got = [counts[i] for i in sorted(counts)]
            avg = float(re.search(r"Average runtime: ([0-9.]+)", out.stdout).group(1))
            gpu = [float(x) for x in re.findall(r"GPU runtime: ([0-9.]+)", out.stdout)]
            results[app] = {"counts": got, "avg_runtime_ms": avg, "gpu_ms_per_run": gpu[-1] if gpu else None,
                            "equal_to_census": sorted(got) == sorted(expect)}
            log(f"{app}: {results[app]}")
            if sorted(got) != sorted(expect):
                raise SystemExit(f"{app} counts {got} differ from the census' {expect}")
    line = base_line("3-vertex motif counts through the unchanged three_motif apps: s, 1 MI355X",
                     results["kautomine_three_motif"]["avg_runtime_ms"] / 1e3, "s", args.steps, 1,
                     results["kautomine_three_motif"]["avg_runtime_ms"], False,
                     {"workload": "chunglu_4.85M_43Mdraws_three_motif_apps", "vertices": g.num_vertices,
                      "edges": g.num_entries // 2})
    line["apps"] = results
    line["census_device_ms"] = census_ms
    print(json.dumps(line), flush=True)
    return 0


def bench_gen(kz, args) -> int:
    """RMAT(scale, edge factor) generation + clean-up + degree orientation on
    the host (OpenMP) and on the GPU; both must give the same CSR."""
    kz.init(0)
    ms = []
    for _ in range(args.warmup + args.steps):
        t0 = time.perf_counter()
        dg = kz.gpu_gen_rmat(args.scale, args.edge_factor, 0.57, 0.19, 0.19, args.seed, orient=True)
        ms.append((time.perf_counter() - t0) * 1e3)
        n, m = dg.num_vertices, dg.num_entries
        last = dg
    gpu_ms = sum(ms[args.warmup:]) / args.steps
    log(f"GPU: {n} vertices, {m} oriented entries, {gpu_ms:.1f} ms")
    host = None
    if not args.no_cpu:
        t0 = time.perf_counter()
        h = kz.gen_rmat(args.scale, args.edge_factor, 0.57, 0.19, 0.19, args.seed).oriented()
        host_s = time.perf_counter() - t0
        same = h.digest() == last.download().digest()
        threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or (os.cpu_count() or 1)
        host = {"seconds": host_s, "threads": threads, "same_digest": same}
        log(f"host: {host}")
        if not same:
            raise SystemExit("GPU and host graphs differ")
    draws = args.edge_factor << args.scale
    line = base_line("graph generation + preprocessing: draws/s, 1 MI355X", draws / (gpu_ms / 1e3), "draws/s",
                     args.steps, args.warmup, gpu_ms, True,
                     {"workload": f"rmat{args.scale}_ef{args.edge_factor}_generate_clean_orient", "draws": draws,
                      "vertices": n, "oriented_edges": m})
    line["host"] = host
    print(json.dumps(line), flush=True)
    return 0


def copy_probe() -> int:
    """Known-size kernels for the HBM byte counters and the measured copy
    bandwidth: 3 copies of 1 GiB (1 GiB read + 1 GiB written each), then 3
    read-only sums of the same 1 GiB (nothing written but the sum), then 20
    timed copies of 4 GiB."""
    import torch

    n = 1 << 28
    a = torch.ones(n, dtype=torch.int32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    for _ in range(3):
        a.sum(dtype=torch.int64)
    torch.cuda.synchronize()
    del a, b
    big = 1 << 30
    x = torch.ones(big, dtype=torch.int32, device="cuda")
    y = torch.empty_like(x)
    y.copy_(x)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(20):
        y.copy_(x)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / 20
    print(json.dumps({"probe_bytes_read": n * 4, "probe_bytes_written": n * 4, "probe_sum_bytes_read": n * 4,
                      "copy_GBps": 2 * 4 * big / (ms / 1e3) / 1e9, "copy_bytes_each": 2 * 4 * big,
                      "copy_ms": ms}), flush=True)
    return 0


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--scale", type=int, default=24)
    ap.add_argument("--edge-factor", type=int, default=16)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--k", type=int, default=4)
    ap.add_argument("--plain-steps", type=int, default=1)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-dgraph", action="store_true", help="skip the .dgraph round trip")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run known-size copies and read-only sums (calibrates the HBM byte counters) "
                         "and time a 4 GiB copy")
    ap.add_argument("--workload", choices=["clique", "census", "batch", "serve", "gen", "fsm", "apps"],
                    default="clique")
    ap.add_argument("--labels", type=int, default=8, help="fsm: number of vertex labels")
    ap.add_argument("--threshold", type=int, default=20000, help="fsm: MNI support threshold")
    ap.add_argument("--host-gen", action="store_true", help="batch/serve: generate on the host, not the GPU")
    ap.add_argument("--gpu-gen", action="store_true",
                    help="clique: build the graph on the GPU (no .dgraph round trip; large scales)")
    ap.add_argument("--graphs", type=int, default=256, help="batch/serve: number of pool graphs")
    ap.add_argument("--check", action="store_true", help="census/batch: also check against the oracle")
    ap.add_argument("--rates", default="4,8,16,32", help="serve: arrival rates (queries/s)")
    ap.add_argument("--seconds", type=float, default=8.0, help="serve: seconds of arrivals per rate")
    ap.add_argument("--budget-gb", type=float, default=64.0, help="serve: HBM budget for prepared executions")
    ap.add_argument("--large-share", type=float, default=0.5, help="serve: share of queries on the large graph")
    args = ap.parse_args()
    if args.copy_probe:
        return copy_probe()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one MI355X: the library has no multi-GPU path (run with --gpus 1)",
              file=sys.stderr)
        return 2

    import openkhuzdulgraph_amd as kz

    if args.workload == "census":
        return bench_census(kz, args)
    if args.workload == "fsm":
        return bench_fsm(kz, args)
    if args.workload == "batch":
        return bench_batch(kz, args)
    if args.workload == "serve":
        return bench_serve(kz, args)
    if args.workload == "gen":
        return bench_gen(kz, args)
    if args.workload == "apps":
        return bench_apps(kz, args)

    return bench_clique(kz, args)


def golden_entry(args, digest: str):
    """The committed exact count of this workload's graph (oracle and, for
    the small graphs, the reference binaries), when one exists for it."""
    path = ROOT / "tests" / "golden" / "config_counts.json"
    if not path.exists() or args.seed != 1:
        return None
    for e in json.loads(path.read_text()).values():
        if e.get("digest") == digest and e.get("oriented"):
            apps = {3: "kautomine_triangle", 4: "kautomine_four_clique", 5: "kautomine_five_clique"}
            c = e.get("counts", {}).get(f"oracle:{apps[args.k]}")
            if c is not None:
                return c
    return None


def host_cpu() -> dict:
    model = None
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count()
    return {"nproc": usable, "cpu_model": model}


def bench_clique(kz, args) -> int:
    from openkhuzdulgraph_amd import workloads as wl

    t0 = time.perf_counter()
    if args.gpu_gen:
        kz.init(0)
        dg = wl.rmat(kz, args.scale, args.edge_factor, args.seed, gpu=True)
        log(f"built on the GPU: {dg.num_vertices} vertices, {dg.num_entries} oriented entries in "
            f"{time.perf_counter() - t0:.1f} s")
        g = dg.download()
    else:
        g = wl.rmat(kz, args.scale, args.edge_factor, args.seed)
        log(f"generated {g} in {time.perf_counter() - t0:.1f} s")
    digest = f"{g.digest():016x}"
    if not args.no_dgraph and not args.gpu_gen:
        with tempfile.TemporaryDirectory() as td:
            path = os.path.join(td, f"rmat{args.scale}.dgraph")
            g.store(path)
            g2 = kz.load_dgraph(path)
            if f"{g2.digest():016x}" != digest:
                raise SystemExit(".dgraph round trip changed the graph")
            g = g2
        log(".dgraph store/load round trip ok")

    kz.init(0)
    dev = kz.device_info()
    if not args.gpu_gen:
        dg = g.to_device()
    work = dg.clique_workload()
    log(f"workload: {work}")

    # persistent form
    for _ in range(args.warmup):
        count = dg.count_cliques(args.k, "persistent")
    dev_ms = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        c = dg.count_cliques(args.k, "persistent")
        dev_ms.append(dg.last_stats["device_ms"])
        if args.warmup == 0:
            count = c
        if c != count:
            raise SystemExit(f"persistent count changed between steps: {c} vs {count}")
    t_pers = (time.perf_counter() - t0) / args.steps
    count = c
    dev_avg_ms = sum(dev_ms) / len(dev_ms)
    log(f"persistent: count={count} {t_pers * 1e3:.2f} ms/step (device {dev_avg_ms:.2f} ms)")
    golden = golden_entry(args, digest)
    if golden is not None and golden != count:
        raise SystemExit(f"count {count} differs from the committed oracle count {golden}")

    # plain level-by-level form
    plain = None
    if args.plain_steps > 0:
        tp = []
        for _ in range(args.plain_steps):
            t1 = time.perf_counter()
            cp = dg.count_cliques(args.k, "plain")
            tp.append(time.perf_counter() - t1)
            if cp != count:
                raise SystemExit(f"plain count {cp} != persistent count {count}")
        st = dg.last_stats
        t_plain = min(tp)
        plain = {"seconds": t_plain, "device_ms": st["device_ms"], "chunks": st["chunks"],
                 "launches": st["launches"], "embeddings": st["embeddings"]}
        log(f"plain: {t_plain:.3f} s")

    B4 = work["B4"] if args.k == 4 else None
    # bytes the vertex-centric kernels must stream: N+(u) once per root and
    # N+(v) once per edge (u, v) (column ids, 4 B each)
    b_vc = 4 * (work["m"] + work["sum_dv"])
    achieved = b_vc / (dev_avg_ms / 1e3) / 1e9
    roofline = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                "kernel": "clique_block_kernel/clique_wave_kernel (persistent bucket graph, one hipGraph launch)",
                "bytes_per_launch": b_vc, "avg_launch_ms": dev_avg_ms}
    traffic_file = ROOT / "profiles" / f"rmat{args.scale}_k{args.k}_persistent_traffic.json"
    if traffic_file.exists():
        tr = json.loads(traffic_file.read_text())
        roofline["traffic"] = tr.get("bytes_per_launch")
        roofline["l2_hit_rate"] = tr.get("l2_hit_rate")
        roofline["measured_copy_GBps"] = tr.get("copy_GBps")

    cpu = None
    if not args.no_cpu:
        threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or (os.cpu_count() or 1)
        cpu = cpu_baseline(g, args.cpu_seconds, threads)
        # the sample's time scaled to the whole count by merge-model bytes
        wb = work["B4"] if args.k == 4 else work["B3"]
        cpu["value"] = cpu["sample_seconds"] * wb / cpu["sample_bytes"] if args.k == 4 else None
        cpu["projection"] = "sample seconds x (merge-model bytes of the whole count / of the sample)"
        cpu.update(host_cpu())
        log(f"cpu baseline: {cpu}")

    line = {
        "metric": METRIC,
        "value": t_pers,
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": t_pers * 1e3,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "uint32",
        "data": "synthetic",
        "config": {
            "workload": f"rmat{args.scale}_ef{args.edge_factor}_k{args.k}_clique_count",
            "generator": "rmat a/b/c/d=0.57/0.19/0.19/0.05, symmetrised, deduped, no loops/isolated, degree-oriented",
            "scale": args.scale, "edge_factor": args.edge_factor, "seed": args.seed, "k": args.k,
            "vertices": dg.num_vertices, "oriented_edges": dg.num_entries, "max_out_degree": dg.max_degree,
            "csr_digest": digest,
        },
        "count": count,
        "golden_count": golden,
        "seconds": t_pers,
        "device_ms": dev_avg_ms,
        "counts_per_s": count / t_pers,
        "edges_per_s": dg.num_entries / t_pers,
        "streamed_GBps": achieved,
        "merge_model_bytes": B4,
        "merge_model_equivalent_GBps": (B4 / t_pers / 1e9) if B4 else None,
        "plain": plain,
        "speedup_vs_plain": (plain["seconds"] / t_pers) if plain else None,
        "roofline": roofline,
        "cpu_baseline": cpu,
        "device": {"arch": dev.get("arch"), "cus": dev.get("compute_units")},
    }
    print(json.dumps(line), flush=True)
    dg.release()
    return 0


if __name__ == "__main__":
    sys.exit(main())
