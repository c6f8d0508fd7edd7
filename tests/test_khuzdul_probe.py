"""Drop-in boundary, CPU side: the reference's generated apps, compiled
unchanged against openkhuzdulgraph_amd/khuzdul/include, are probed into
plans (KZ_PROBE_ONLY=1, KZ_PLAN_DUMP) and the plans, run by the oracle's
plan interpreter, give the counts the REFERENCE binaries printed on the
golden fixtures.  A plan that mis-states one predicate bit or operand fails
here."""

import ctypes as C
import json
import os
import subprocess
from pathlib import Path

import pytest

import openkhuzdulgraph_amd as kz
import oracle
from openkhuzdulgraph_amd import _ffi

ROOT = Path(__file__).resolve().parent.parent
APPS_DIR = ROOT / "openkhuzdulgraph_amd" / "khuzdul" / "_apps"
GOLDEN = ROOT / "tests" / "golden"
# multiplicities the apps divide by (kautomine/three_motif.cc:44; others 1)
MULT = {"kautomine_three_motif": [1, 2]}


def app_plans(app: str, graph: Path, tmp_path: Path):
    dump = tmp_path / f"{app}.plans"
    if dump.exists():
        dump.unlink()
    env = dict(os.environ, KZ_PROBE_ONLY="1", KZ_PLAN_DUMP=str(dump))
    subprocess.run([str(APPS_DIR / app), str(graph), "1", "1"], env=env, check=True, capture_output=True,
                   timeout=120)
    raw = dump.read_bytes()
    sz = C.sizeof(_ffi.kz_plan_desc)
    assert len(raw) % sz == 0
    return [_ffi.kz_plan_desc.from_buffer_copy(raw[i:i + sz]) for i in range(0, len(raw), sz)]


def built_apps():
    if not APPS_DIR.is_dir():
        return []
    return sorted(p.name for p in APPS_DIR.iterdir() if p.is_file())


@pytest.mark.skipif(not built_apps(), reason="apps are built from the reference sources (make -C openkhuzdulgraph_amd/khuzdul apps)")
@pytest.mark.parametrize("app", sorted(oracle.APPS))
def test_probed_plans_reproduce_reference_counts(app, tmp_path):
    golden = json.loads((GOLDEN / "golden_counts.json").read_text())
    for name, e in golden.items():
        want = e["counts"][app]
        if not want:
            continue
        g = kz.load_dgraph(str(GOLDEN / e["file"]))
        plans = app_plans(app, GOLDEN / e["file"], tmp_path)
        got = []
        for p in plans:
            assert kz._ffi.lib().kz_plan_validate(C.byref(p)) == 0
            got.extend(oracle.run_plan(p, g.rowptr, g.col))
        mult = MULT.get(app, [1] * len(got))
        got = [x // m for x, m in zip(got, mult)]
        assert got == want, (app, name, got, want)


@pytest.mark.skipif(not built_apps(), reason="apps not built")
def test_clique_apps_probe_to_the_builtin_clique_plan(tmp_path):
    g = GOLDEN / "dense_40_600_s4.dgraph"
    for fam in ("kautomine", "kgraphpi"):
        for k, name in ((3, "triangle"), (4, "four_clique"), (5, "five_clique")):
            (p,) = app_plans(f"{fam}_{name}", g, tmp_path)
            ref = kz.clique_plan(k)
            assert p.num_step_lists == ref.num_step_lists
            assert bytes(p.type_table) == bytes(ref.type_table)
            for i in range(p.num_step_lists):
                a, b = p.lists[i], ref.lists[i]
                assert a.num_ops == b.num_ops and a.num_outs == b.num_outs
                assert bytes(a.ops) == bytes(b.ops)
                for o in range(a.num_outs):
                    assert (a.outs[o].kind, a.outs[o].pred, bytes(a.outs[o].src)) == \
                        (b.outs[o].kind, b.outs[o].pred, bytes(b.outs[o].src))


@pytest.mark.skipif(not built_apps(), reason="apps not built")
def test_fused_three_motif_has_three_order_types(tmp_path):
    (p,) = app_plans("kautomine_three_motif", GOLDEN / "dense_40_600_s4.dgraph", tmp_path)
    assert p.num_aggregators == 2
    types = [c for c in range(256) if p.type_table[2][c] != 255]
    assert sorted(types) == [0, 1, 2]  # tie, v1 < v0, v0 < v1
