"""Every measured workload at full size, bit-exact on the GPU.

The graphs are built in HBM by the GPU generators, their CSR digests must
equal the host generators' (recorded in tests/golden/config_counts.json by
tests/golden/make_config_golden.py), and their counts must equal the
committed counts: the reference binaries' (one MPI rank) where the graph is
small enough to run them, else the oracle's (the C++ restatement of the
kautomine apps, pinned against the reference binaries on the fixtures).
kautomine/triangle.cc:58-121 and kautomine/four_clique.cc:58-154 are the
schedules counted."""

import json
import os
import re
import subprocess
from pathlib import Path

import pytest

pytestmark = pytest.mark.gpu

ROOT = Path(__file__).resolve().parent.parent
APPS_DIR = ROOT / "openkhuzdulgraph_amd" / "khuzdul" / "_apps"
GOLDEN = json.loads((ROOT / "tests" / "golden" / "config_counts.json").read_text())
PAT = re.compile(r"Pattern count (\d+): (\d+)")


def _entry(name):
    if name not in GOLDEN:
        pytest.skip(f"no committed count for {name}")
    return GOLDEN[name]


def _check_graph(dg, e):
    h = dg.download()
    assert (h.num_vertices, h.num_entries, h.max_degree, h.oriented_flag) == \
        (e["num_vertices"], e["num_entries"], e["max_degree"], e["oriented"])
    assert f"{h.digest():016x}" == e["digest"]
    return h


def test_c1_er_triangles_equal_the_reference_binaries(kzdev, tmp_path):
    from openkhuzdulgraph_amd import workloads as wl
    e = _entry("c1_er_2p16_2p20_triangle")
    want = e["counts"]["reference:kautomine_triangle"]
    assert want == e["counts"]["reference:kgraphpi_triangle"] == e["counts"]["oracle:kautomine_triangle"]
    print("[case] C1 ER 2^16/2^20 symmetric", flush=True)
    dg = wl.er_2p16_2p20(kzdev, gpu=True)
    h = _check_graph(dg, e)
    assert dg.count_cliques(3, "persistent") == want
    assert dg.count_cliques(3, "plain") == want
    dg.release()
    if APPS_DIR.is_dir():
        path = tmp_path / "er.dgraph"
        h.store(str(path))
        for app in ("kautomine_triangle", "kgraphpi_triangle"):
            print(f"[case] C1 app={app}", flush=True)
            out = subprocess.run([str(APPS_DIR / app), str(path), "1", "1"], env=dict(os.environ),
                                 capture_output=True, text=True, timeout=120)
            assert out.returncode == 0, out.stderr[-2000:]
            got = {int(m.group(1)): int(m.group(2)) for m in PAT.finditer(out.stdout)}
            assert got.get(0) == want, out.stdout[-1500:]


def test_c2_rmat22_cliques_equal_the_oracle(kzdev):
    from openkhuzdulgraph_amd import workloads as wl
    e = _entry("c2_rmat22_ef16")
    print("[case] C2 RMAT-22/16 oriented", flush=True)
    dg = wl.rmat(kzdev, 22, 16, gpu=True)
    _check_graph(dg, e)
    assert dg.count_cliques(3, "persistent") == e["counts"]["oracle:kautomine_triangle"]
    assert dg.count_cliques(4, "persistent") == e["counts"]["oracle:kautomine_four_clique"]
    print("[case] C2 plain 4-clique", flush=True)
    assert dg.count_cliques(4, "plain") == e["counts"]["oracle:kautomine_four_clique"]
    dg.release()


def test_rmat24_four_cliques_equal_the_oracle(kzdev):
    from openkhuzdulgraph_amd import workloads as wl
    e = _entry("target_rmat24_ef16")
    print("[case] RMAT-24/16 oriented 4-clique", flush=True)
    dg = wl.rmat(kzdev, 24, 16, gpu=True)
    _check_graph(dg, e)
    assert dg.count_cliques(4, "persistent") == e["counts"]["oracle:kautomine_four_clique"]
    dg.release()


def test_c3_census_equals_the_oracle(kzdev):
    from openkhuzdulgraph_amd import workloads as wl
    e = _entry("c3_chunglu_census")
    print("[case] C3 Chung-Lu 4.85M census", flush=True)
    dg = wl.chunglu_census(kzdev, gpu=True)
    _check_graph(dg, e)
    assert dg.motif_census() == e["census"]
    dg.release()


def test_c4_pool_subset_batched_equals_the_oracle(kzdev):
    from openkhuzdulgraph_amd import workloads as wl
    e = _entry("c4_pool_subset")
    idx = sorted(int(i) for i in e["graphs"])
    dgs = []
    for i in idx:
        print(f"[case] C4 pool graph {i}", flush=True)
        dg = wl.pool_graph(kzdev, i, gpu=True, orient=True)
        _check_graph(dg, e["graphs"][str(i)])
        dgs.append(dg)
    b = kzdev.Batch(dgs, [3, 4, 5])
    got = b.run()
    b.release()
    for r, i in enumerate(idx):
        want = e["graphs"][str(i)]["k345"]
        assert [int(x) for x, w in zip(got[r], want) if w is not None] == [w for w in want if w is not None], i
    for dg in dgs:
        dg.release()


def test_c5_rmat26_ef32_triangles_equal_the_oracle(kzdev):
    from openkhuzdulgraph_amd import workloads as wl
    e = _entry("c5_rmat26_ef32")
    print("[case] C5 RMAT-26/32 oriented triangles", flush=True)
    dg = wl.rmat(kzdev, 26, 32, gpu=True)
    _check_graph(dg, e)
    assert dg.count_cliques(3, "persistent") == e["counts"]["oracle:kautomine_triangle"]
    dg.release()
