"""Writes the golden fixtures under tests/golden/.

Each fixture is a small seeded graph stored as .dgraph (the reference's own
file format) plus the counts printed by the REFERENCE apps themselves: the
binaries built by oracle/Makefile (`make -C oracle ref`) from the
OpenKhuzdulGraph sources, run as one MPI rank.  The reference holds no
recorded counts of its own, so these vectors are produced here and
committed; the script is re-runnable wherever the reference tree exists.

    python tests/golden/make_golden.py          # clique / motif apps
    python tests/golden/make_golden.py --fsm    # kautomine three_fsm
"""

from __future__ import annotations

import json
import re
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent.parent
sys.path.insert(0, str(ROOT))

import openkhuzdulgraph_amd as kz  # noqa: E402
import oracle  # noqa: E402

# name -> builder; sizes keep the reference's per-level buffers small
FIXTURES = {
    "er_300_3000_s1": lambda: kz.gen_erdos_renyi(300, 3000, 1),
    "er_300_3000_s1_oriented": lambda: kz.gen_erdos_renyi(300, 3000, 1).oriented(),
    "rmat10_ef8_s2": lambda: kz.gen_rmat(10, 8, seed=2),
    "rmat10_ef8_s2_oriented": lambda: kz.gen_rmat(10, 8, seed=2).oriented(),
    "chunglu_2000_s3": lambda: kz.gen_chung_lu(2000, 12000, 2.3, 200, 3),
    "dense_40_600_s4": lambda: kz.gen_erdos_renyi(40, 600, 4),
}

PAT = re.compile(r"Pattern count (\d+): (\d+)")


def run_reference(app: str, path: Path, threads: int = 2) -> list[int]:
    exe = oracle.reference_binary(app)
    if not exe.exists():
        raise SystemExit(f"{exe} missing: run `make -C oracle ref` where the reference tree exists")
    out = subprocess.run([str(exe), str(path), str(threads), "1"], capture_output=True, text=True,
                         timeout=600, cwd="/tmp").stdout
    counts: dict[int, int] = {}
    for m in PAT.finditer(out):
        counts[int(m.group(1))] = int(m.group(2))
    return [counts[i] for i in sorted(counts)]


def main() -> None:
    golden = {}
    for name, make in FIXTURES.items():
        g = make()
        path = HERE / f"{name}.dgraph"
        g.store(str(path))
        entry = {"file": path.name, "num_vertices": g.num_vertices, "num_entries": g.num_entries,
                 "oriented": g.oriented_flag, "digest": f"{g.digest():016x}", "counts": {}}
        for app in oracle.APPS:
            entry["counts"][app] = run_reference(app, path)
        golden[name] = entry
        print(name, entry["counts"], flush=True)
    (HERE / "golden_counts.json").write_text(json.dumps(golden, indent=1, sort_keys=True) + "\n")


# labelled FSM fixtures: name -> (builder, number of labels, label seed,
# support thresholds); labels are drawn uniformly and every label occurs
FSM_FIXTURES = {
    "er_300_3000_s1_l4": (lambda: kz.gen_erdos_renyi(300, 3000, 1), 4, 7, (20, 60)),
    "chunglu_2000_s3_l6": (lambda: kz.gen_chung_lu(2000, 12000, 2.3, 200, 3), 6, 11, (30, 120)),
    "rmat10_ef8_s2_l3": (lambda: kz.gen_rmat(10, 8, seed=2), 3, 5, (50,)),
    "dense_40_600_s4_l2": (lambda: kz.gen_erdos_renyi(40, 600, 4), 2, 3, (5, 19)),
}
FSM_PAT = re.compile(r"Number of frequent ([a-z\- ]+) patterns: (\d+), support sum[a-z ]*: (\d+)")


def fsm_labels(n: int, num_labels: int, seed: int):
    import numpy as np
    return np.random.default_rng(seed).integers(0, num_labels, n).astype(np.uint32)


def run_reference_fsm(path: Path, threshold: int, threads: int = 2) -> dict:
    exe = oracle.reference_binary("kautomine_three_fsm")
    if not exe.exists():
        raise SystemExit(f"{exe} missing: run `make -C oracle ref` where the reference tree exists")
    out = subprocess.run([str(exe), str(path), str(threshold), str(threads)], capture_output=True, text=True,
                         timeout=600, cwd="/tmp").stdout
    rows = FSM_PAT.findall(out)
    if len(rows) != 6:
        raise SystemExit(f"unexpected three_fsm output:\n{out}")
    return {"counts": [int(c) for _, c, _ in rows], "sums": [int(x) for _, _, x in rows]}


def fsm_main() -> None:
    golden = {}
    for name, (make, nl, seed, thresholds) in FSM_FIXTURES.items():
        g = make()
        g.set_labels(fsm_labels(g.num_vertices, nl, seed))
        path = HERE / f"{name}.dgraph"
        g.store(str(path))
        entry = {"file": path.name, "num_vertices": g.num_vertices, "num_entries": g.num_entries,
                 "num_labels": nl, "digest": f"{g.digest():016x}", "three_fsm": {}}
        for t in thresholds:
            entry["three_fsm"][str(t)] = run_reference_fsm(path, t)
        golden[name] = entry
        print(name, entry["three_fsm"], flush=True)
    (HERE / "golden_fsm.json").write_text(json.dumps(golden, indent=1, sort_keys=True) + "\n")


if __name__ == "__main__":
    if sys.argv[1:] == ["--fsm"]:
        fsm_main()
    else:
        main()
