"""VertexSet::intersect_and_count, bounded_intersect_with and
bounded_subtracted_by inside apply() (src/vertex_set.cc:88-184 of the
reference).  No generated app calls them, so three example apps of this
package exercise them (openkhuzdulgraph_amd/khuzdul/examples): triangles by
per-edge intersect_and_count, 4-cliques whose symmetry breaking is done by
bounded intersections (one of them on a cached prefix), and open/closed
wedges by a bounded subtraction.  CPU side: each app is probed into plans
(KZ_PROBE_ONLY=1) that the oracle's plan interpreter runs; the counts must
equal the reference binaries' triangle and 4-clique counts on the golden
fixtures and the closed form open = sum_v C(d_<(v), 2) - triangles, with
d_<(v) the neighbours of v below v."""

import ctypes as C
import json
import os
import subprocess
from pathlib import Path

import numpy as np
import pytest

import openkhuzdulgraph_amd as kz
import oracle
from openkhuzdulgraph_amd import _ffi

ROOT = Path(__file__).resolve().parent.parent
EX_DIR = ROOT / "openkhuzdulgraph_amd" / "khuzdul" / "_examples"
GOLDEN = ROOT / "tests" / "golden"
# symmetric fixtures (the examples match vertices in decreasing id order)
FIXTURES = ["dense_40_600_s4", "er_300_3000_s1", "chunglu_2000_s3", "rmat10_ef8_s2"]


def expected(name: str):
    """Counts each example prints on a golden fixture."""
    e = json.loads((GOLDEN / "golden_counts.json").read_text())[name]
    tri = e["counts"]["kautomine_triangle"][0]
    k4 = e["counts"]["kautomine_four_clique"][0]
    g = kz.load_dgraph(str(GOLDEN / e["file"]))
    rp, col = np.asarray(g.rowptr, dtype=np.int64), np.asarray(g.col, dtype=np.int64)
    src = np.repeat(np.arange(len(rp) - 1), np.diff(rp))
    dl = np.bincount(src[col < src], minlength=len(rp) - 1).astype(np.int64)
    wedges = int((dl * (dl - 1) // 2).sum())
    return {
        "example_edge_triangles": ([tri], [3 * tri]),
        "example_bounded_four_clique": ([k4], [k4]),
        "example_bounded_wedges": ([wedges - tri, tri], [wedges - tri, 3 * tri]),
    }


def probed_plans(app: str, graph: Path, tmp_path: Path):
    dump = tmp_path / f"{app}.plans"
    if dump.exists():
        dump.unlink()
    env = dict(os.environ, KZ_PROBE_ONLY="1", KZ_PLAN_DUMP=str(dump))
    subprocess.run([str(EX_DIR / app), str(graph)], env=env, check=True, capture_output=True, timeout=120)
    raw = dump.read_bytes()
    sz = C.sizeof(_ffi.kz_plan_desc)
    assert len(raw) % sz == 0
    return [_ffi.kz_plan_desc.from_buffer_copy(raw[i:i + sz]) for i in range(0, len(raw), sz)]


@pytest.mark.skipif(not EX_DIR.is_dir(), reason="examples not built (make -C openkhuzdulgraph_amd/khuzdul examples)")
@pytest.mark.parametrize("app", ["example_edge_triangles", "example_bounded_four_clique", "example_bounded_wedges"])
def test_bounded_ops_probe_to_exact_plans(app, tmp_path):
    for name in FIXTURES:
        _, want_raw = expected(name)[app]
        g = kz.load_dgraph(str(GOLDEN / f"{name}.dgraph"))
        got = []
        for p in probed_plans(app, GOLDEN / f"{name}.dgraph", tmp_path):
            assert kz._ffi.lib().kz_plan_validate(C.byref(p)) == 0
            got.extend(oracle.run_plan(p, g.rowptr, g.col))
        assert got == want_raw, (app, name, got, want_raw)


@pytest.mark.skipif(not EX_DIR.is_dir(), reason="examples not built")
def test_bound_becomes_a_predicate_below_the_matched_vertex(tmp_path):
    """bounded_subtracted_by(N(v1), ub = v1) at (v0 > v1): the subtraction
    is a plain N(m0) \\ N(m1) and its count keeps position class 0 only
    (below m1, the lower of the two matched values)."""
    (p,) = probed_plans("example_bounded_wedges", GOLDEN / "dense_40_600_s4.dgraph", tmp_path)
    code = 1  # ranks (1, 0): v1 < v0
    L = p.lists[p.type_table[2][code]]
    assert L.num_ops == 2
    assert [L.ops[j].op for j in range(2)] == [2, 1]  # subtract, then intersect
    preds = sorted((L.outs[o].aggregator, L.outs[o].pred) for o in range(L.num_outs))
    assert preds == [(0, 0b1), (1, 0b11111)]
