"""Writes tests/golden/config_counts.json: exact counts of the measured
workloads' own graphs at full size, for the GPU tests to compare against.

For each workload the graph is generated on the host (the GPU builds the
same CSR, same digest), its size and CSR digest are recorded, and its counts
come from the oracle (oracle/kz_oracle.cpp, the C++ restatement of the
kautomine/kgraphpi apps pinned against the reference binaries on the small
fixtures) and, where the reference apps built by `make -C oracle ref` are
present and the graph is small enough, from those binaries run as one MPI
rank on the graph's .dgraph file.  Sizes and times of the runs that made the
file are recorded beside the counts.

    python tests/golden/make_config_golden.py [--only NAME ...] [--threads T]
           [--reference NAME ...]

Entries are merged into the existing file, so the slow ones (the scale-24
4-clique count takes about an hour on 8 cores) can run on their own.
"""

from __future__ import annotations

import argparse
import json
import os
import re
import subprocess
import sys
import tempfile
import time
from pathlib import Path

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent.parent
sys.path.insert(0, str(ROOT))

import openkhuzdulgraph_amd as kz  # noqa: E402
from openkhuzdulgraph_amd import workloads as wl  # noqa: E402
import oracle  # noqa: E402

OUT = HERE / "config_counts.json"
PAT = re.compile(r"Pattern count (\d+): (\d+)")

# pool graphs whose clique counts are recorded: every 256th of the 4096
# (Erdos-Renyi, even) and the RMAT graph after each (odd)
POOL_SUBSET = sorted(i + j for i in range(0, wl.POOL_SIZE, 256) for j in (0, 1))


def graph_entry(g) -> dict:
    return {"num_vertices": g.num_vertices, "num_entries": g.num_entries, "max_degree": g.max_degree,
            "oriented": g.oriented_flag, "digest": f"{g.digest():016x}"}


def run_reference(app: str, g, threads: int) -> tuple[list[int], float]:
    """Counts printed by a reference app run as one MPI rank; the reference
    engine takes 1-6 compute threads on an 8-core host
    (distributed_application.cc:80)."""
    exe = oracle.reference_binary(app)
    if not exe.exists():
        raise SystemExit(f"{exe} missing: run `make -C oracle ref` where the reference tree exists")
    with tempfile.TemporaryDirectory() as td:
        path = os.path.join(td, "g.dgraph")
        g.store(path)
        t0 = time.perf_counter()
        out = subprocess.run([str(exe), path, str(threads), "1"], capture_output=True, text=True,
                             cwd="/tmp", check=True).stdout
        dt = time.perf_counter() - t0
    counts = {int(m.group(1)): int(m.group(2)) for m in PAT.finditer(out)}
    return [counts[i] for i in sorted(counts)], dt


def oracle_app(app: str, g, threads: int) -> tuple[int, float]:
    t0 = time.perf_counter()
    _, cnt = oracle.run_app(app, g.rowptr, g.col, threads=threads)
    return cnt[0], time.perf_counter() - t0


def c1(threads: int, reference: bool) -> dict:
    g = wl.er_2p16_2p20(kz)
    e = {"workload": "Erdos-Renyi n=2^16, m=2^20 draws, seed 1, symmetric", **graph_entry(g), "counts": {}}
    for app in ("kautomine_triangle", "kgraphpi_triangle"):
        c, dt = oracle_app(app, g, threads)
        e["counts"][f"oracle:{app}"] = c
        e[f"oracle_seconds:{app}"] = round(dt, 2)
    if reference:
        for app in ("kautomine_triangle", "kgraphpi_triangle", "kautomine_four_clique"):
            c, dt = run_reference(app, g, min(threads, 6))
            e["counts"][f"reference:{app}"] = c[0]
            e[f"reference_seconds:{app}"] = round(dt, 2)
    return e


def rmat_entry(scale: int, ef: int, ks: tuple, threads: int, reference: bool) -> dict:
    g = wl.rmat(kz, scale, ef)
    e = {"workload": f"RMAT scale {scale}, edge factor {ef}, a/b/c/d .57/.19/.19/.05, seed 1, degree-oriented",
         **graph_entry(g), "counts": {}}
    apps = {3: "kautomine_triangle", 4: "kautomine_four_clique", 5: "kautomine_five_clique"}
    for k in ks:
        c, dt = oracle_app(apps[k], g, threads)
        e["counts"][f"oracle:{apps[k]}"] = c
        e[f"oracle_seconds:{apps[k]}"] = round(dt, 1)
        print(f"  rmat{scale} k={k}: {c} in {dt:.1f} s", flush=True)
    if reference:
        for k in ks:
            c, dt = run_reference(apps[k], g, 6)
            e["counts"][f"reference:{apps[k]}"] = c[0]
            e[f"reference_seconds:{apps[k]}"] = round(dt, 1)
    e["threads"] = threads
    return e


def c3(threads: int, reference: bool) -> dict:
    g = wl.chunglu_census(kz)
    t0 = time.perf_counter()
    census = oracle.motif_census(g.rowptr, g.col, threads=threads)
    e = {"workload": "Chung-Lu n=4.85M, 43M draws, exponent 2.3, weight cap 20000, seed 1, symmetric",
         **graph_entry(g), "census_patterns": ["wedge", "triangle", "3-path", "3-star", "4-cycle",
                                               "tailed-triangle", "diamond", "4-clique"],
         "census": census, "oracle_seconds": round(time.perf_counter() - t0, 1), "threads": threads}
    return e


# 5-cliques are recorded for pool graphs of at most this many draws (the
# oracle's enumeration of the densest RMAT graphs' 5-cliques takes hours)
POOL_K5_MAX_DRAWS = 1 << 21


def c4(threads: int, reference: bool) -> dict:
    e = {"workload": f"pool graphs i and i + 1 for i = 0, 256, ..., {wl.POOL_SIZE - 256} of {wl.POOL_SIZE} "
                     "(Erdos-Renyi and RMAT), degree-oriented; triangle, 4- and 5-clique (5-cliques of graphs "
                     f"of at most {POOL_K5_MAX_DRAWS} draws; null otherwise)", "graphs": {}}
    apps = ("kautomine_triangle", "kautomine_four_clique", "kautomine_five_clique")
    for i in POOL_SUBSET:
        g = wl.pool_graph(kz, i, orient=True)
        kind, a, b, _ = wl.pool_params(i)
        draws = b if kind == "er" else b << a
        t0 = time.perf_counter()
        cnt = [oracle.run_app(app, g.rowptr, g.col, threads=threads)[1][0]
               if app != apps[2] or draws <= POOL_K5_MAX_DRAWS else None for app in apps]
        e["graphs"][str(i)] = {"params": list(wl.pool_params(i)), **graph_entry(g), "k345": cnt,
                               "oracle_seconds": round(time.perf_counter() - t0, 1)}
        print(f"  pool {i}: {cnt}", flush=True)
    return e


def sym_rmat16(threads: int, reference: bool) -> dict:
    """Symmetric RMAT-16 (no orientation): its low-id hubs keep thousands of
    higher-id neighbours, so a by-id forward part would exceed the bitmap
    kernels' degree limit; the clique apps must still count it."""
    g = wl.rmat(kz, 16, 16, orient=False)
    rp, col = g.rowptr, g.col
    fwd_by_id = int(max((col[rp[v]:rp[v + 1]] > v).sum() for v in range(g.num_vertices)))
    e = {"workload": "RMAT scale 16, edge factor 16, a/b/c/d .57/.19/.19/.05, seed 1, symmetric",
         **graph_entry(g), "max_forward_degree_by_id": fwd_by_id, "counts": {}}
    apps = ("kautomine_triangle", "kgraphpi_triangle", "kautomine_four_clique", "kgraphpi_four_clique")
    for app in apps:
        c, dt = oracle_app(app, g, threads)
        e["counts"][f"oracle:{app}"] = c
    if reference:
        for app in apps:
            c, dt = run_reference(app, g, min(threads, 6))
            e["counts"][f"reference:{app}"] = c[0]
            e[f"reference_seconds:{app}"] = round(dt, 2)
    return e


CONFIGS = {
    "sym_rmat16_ef16": sym_rmat16,
    "c1_er_2p16_2p20_triangle": c1,
    "c2_rmat22_ef16": lambda t, r: rmat_entry(22, 16, (3, 4), t, r),
    "target_rmat24_ef16": lambda t, r: rmat_entry(24, 16, (4,), t, r),
    "c3_chunglu_census": c3,
    "c4_pool_subset": c4,
    "c5_rmat26_ef32": lambda t, r: rmat_entry(26, 32, (3,), t, r),
}


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--only", nargs="*", default=None, choices=list(CONFIGS))
    ap.add_argument("--reference", nargs="*", default=["c1_er_2p16_2p20_triangle", "sym_rmat16_ef16"],
                    help="workloads whose graph also goes through the reference binaries")
    ap.add_argument("--threads", type=int, default=os.cpu_count() or 1)
    args = ap.parse_args()
    data = json.loads(OUT.read_text()) if OUT.exists() else {}
    for name in args.only or list(CONFIGS):
        t0 = time.perf_counter()
        print(f"{name} ...", flush=True)
        e = CONFIGS[name](args.threads, name in args.reference)
        e["host"] = {"nproc": os.cpu_count(), "seconds": round(time.perf_counter() - t0, 1)}
        data = json.loads(OUT.read_text()) if OUT.exists() else {}
        data[name] = e
        OUT.write_text(json.dumps(data, indent=1, sort_keys=True) + "\n")
        print(f"{name}: done in {e['host']['seconds']} s", flush=True)


if __name__ == "__main__":
    main()
