"""Benchmark: exact 4-clique count on a seeded RMAT graph, one MI355X.

Workload (default): RMAT scale 22, edge factor 16, a/b/c/d = .57/.19/.19/.05,
seed 1; symmetrised, deduplicated, self-loops and isolated vertices dropped,
degree-oriented; uint32 CSR written to and loaded back from a .dgraph file.
The graph is resident in HBM before timing starts.

A step is one full exact count of the graph's 4-cliques.  The persistent
form (one hipGraph launch of the degree-bucketed kernels) is the headline;
the plain level-by-level form is timed beside it.  Throughput `value` is the
merge-model algorithmic bytes B4 of the count divided by the persistent
form's seconds (GB/s).  The CPU baseline is the oracle's restatement of the
kautomine four_clique schedule on a bounded sample of roots, on this host's
cores.

    python bench.py [--steps K] [--warmup W] [--scale 22] [--plain-steps 1]
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
        "value": b["B4"] / dt / 1e9,
        "unit": "GB/s",
        "cores": threads,
        "kind": "port",
        "sample": f"every {stride}-th root vertex ({len(roots)} of {n}) of the same graph, "
                  f"{cnt[0]} 4-cliques, {b['B4']} merge-model bytes, {dt:.2f} s",
        "seconds": dt,
    }


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--scale", type=int, default=22)
    ap.add_argument("--edge-factor", type=int, default=16)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--k", type=int, default=4)
    ap.add_argument("--plain-steps", type=int, default=1)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-dgraph", action="store_true", help="skip the .dgraph round trip")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run a 1 GiB device-to-device copy (calibrates HBM byte counters)")
    args = ap.parse_args()
    if args.copy_probe:
        import torch

        a = torch.ones(1 << 28, dtype=torch.int32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes_read": a.numel() * 4, "copy_probe_bytes_written": b.numel() * 4}))
        return 0
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one MI355X: this build has no multi-GPU path (run with --gpus 1)",
              file=sys.stderr)
        return 2

    import openkhuzdulgraph_amd as kz

    t0 = time.perf_counter()
    sym = kz.gen_rmat(args.scale, args.edge_factor, 0.57, 0.19, 0.19, args.seed)
    g = sym.oriented()
    del sym
    log(f"generated {g} in {time.perf_counter() - t0:.1f} s")
    digest = g.digest()
    if not args.no_dgraph:
        with tempfile.TemporaryDirectory() as td:
            path = os.path.join(td, f"rmat{args.scale}.dgraph")
            g.store(path)
            g2 = kz.load_dgraph(path)
            if g2.digest() != digest:
                raise SystemExit(".dgraph round trip changed the graph")
            g = g2
        log(".dgraph store/load round trip ok")

    kz.init(0)
    dev = kz.device_info()
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
    log(f"persistent: count={count} {t_pers * 1e3:.2f} ms/step (device {sum(dev_ms) / len(dev_ms):.2f} ms)")

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
                 "launches": st["launches"], "embeddings": st["embeddings"],
                 "GB/s": work["B4"] / t_plain / 1e9}
        log(f"plain: {t_plain:.3f} s")

    B4 = work["B4"] if args.k == 4 else None
    dev_avg_ms = sum(dev_ms) / len(dev_ms)
    # bytes the vertex-centric kernel must stream: N+(u) once per root and
    # N+(v) once per edge (u, v) (column ids, 4 B each)
    b_vc = 4 * (work["m"] + work["sum_dv"])
    achieved = b_vc / (dev_avg_ms / 1e3) / 1e9
    roofline = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                "kernel": "clique_block_kernel/clique_wave_kernel (persistent bucket graph)",
                "bytes_per_launch": b_vc, "avg_launch_ms": dev_avg_ms}
    traffic_file = ROOT / "profiles" / f"rmat{args.scale}_k{args.k}_persistent_traffic.json"
    if traffic_file.exists():
        roofline["traffic"] = json.loads(traffic_file.read_text()).get("bytes_per_launch")

    cpu = None
    if not args.no_cpu:
        threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or (os.cpu_count() or 1)
        cpu = cpu_baseline(g, args.cpu_seconds, threads)
        log(f"cpu baseline: {cpu}")

    value = (B4 / t_pers / 1e9) if B4 else None
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": t_pers * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "uint32",
        "data": "synthetic",
        "config": {
            "workload": f"rmat{args.scale}_ef{args.edge_factor}_k{args.k}_clique_count",
            "generator": "rmat a/b/c/d=0.57/0.19/0.19/0.05, symmetrised, deduped, no loops/isolated, degree-oriented",
            "scale": args.scale, "edge_factor": args.edge_factor, "seed": args.seed, "k": args.k,
            "vertices": g.num_vertices, "oriented_edges": g.num_entries, "max_out_degree": g.max_degree,
            "csr_digest": f"{digest:016x}",
        },
        "count": count,
        "seconds": t_pers,
        "counts_per_s": count / t_pers,
        "edges_per_s": g.num_entries / t_pers,
        "merge_model_bytes": B4,
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
