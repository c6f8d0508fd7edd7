"""Drop-in boundary on the GPU: the reference's generated apps, compiled
unchanged against libkhuzdul_amd.so, print on the golden fixtures exactly
the counts the reference binaries printed; the plain plan executor matches
the oracle's plan interpreter on every probed plan, including the fused
3-motif (two aggregators, three order types)."""

import ctypes as C
import json
import os
import re
import subprocess
from pathlib import Path

import pytest

import oracle
from openkhuzdulgraph_amd import _ffi

pytestmark = pytest.mark.gpu

ROOT = Path(__file__).resolve().parent.parent
APPS_DIR = ROOT / "openkhuzdulgraph_amd" / "khuzdul" / "_apps"
GOLDEN = ROOT / "tests" / "golden"
PAT = re.compile(r"Pattern count (\d+): (\d+)")


def _apps():
    return sorted(p.name for p in APPS_DIR.iterdir()) if APPS_DIR.is_dir() else []


@pytest.mark.skipif(not _apps(), reason="app binaries were not built from the reference sources")
@pytest.mark.parametrize("method", ["auto", "plain"])
def test_unchanged_apps_print_reference_counts(method):
    golden = json.loads((GOLDEN / "golden_counts.json").read_text())
    for app in sorted(oracle.APPS):
        for name, e in golden.items():
            want = e["counts"][app]
            if not want:
                continue
            print(f"[case] app={app} graph={name} method={method}", flush=True)
            env = dict(os.environ)
            if method == "plain":
                env["KZ_METHOD"] = "plain"
            out = subprocess.run([str(APPS_DIR / app), str(GOLDEN / e["file"]), "1", "1"], env=env,
                                 capture_output=True, text=True, timeout=120)
            assert out.returncode == 0, out.stderr[-2000:]
            got = {}
            for m in PAT.finditer(out.stdout):
                got[int(m.group(1))] = int(m.group(2))
            assert [got[i] for i in sorted(got)] == want, (app, name, out.stdout[-1500:])


@pytest.mark.skipif(not _apps(), reason="app binaries were not built from the reference sources")
@pytest.mark.parametrize("flat", ["1", "0"])
def test_plan_executor_matches_oracle_on_probed_plans(kzdev, tmp_path, monkeypatch, flat):
    """Both paths of the executor's counting sizes: sibling runs streamed
    as one (flat, the default) and one embedding per wave (KZ_PLAN_FLAT=0).
    The Chung-Lu graph with exponent 2.0 has 14 vertices of degree above
    1,024, whose lists are probed by binary search instead of the LDS hash."""
    kz = kzdev
    monkeypatch.setenv("KZ_PLAN_FLAT", flat)
    graphs = [kz.gen_rmat(11, 16, seed=8), kz.gen_chung_lu(3000, 20000, 2.3, 300, 4),
              kz.gen_rmat(11, 16, seed=8).oriented(), kz.gen_chung_lu(20000, 150000, 2.0, 4000, 11)]
    for app in sorted(oracle.APPS):
        dump = tmp_path / f"{app}.plans"
        env = dict(os.environ, KZ_PROBE_ONLY="1", KZ_PLAN_DUMP=str(dump))
        subprocess.run([str(APPS_DIR / app), str(GOLDEN / "dense_40_600_s4.dgraph"), "1", "1"], env=env, check=True,
                       capture_output=True, timeout=120)
        raw = dump.read_bytes()
        sz = C.sizeof(_ffi.kz_plan_desc)
        plans = [_ffi.kz_plan_desc.from_buffer_copy(raw[i:i + sz]) for i in range(0, len(raw), sz)]
        for g in graphs:
            dg = g.to_device()
            for p in plans:
                print(f"[case] app={app} graph={g} flat={flat}", flush=True)
                want = oracle.run_plan(p, g.rowptr, g.col)
                assert dg.count_plan(p) == want, (app, str(g))
            dg.release()


EX_DIR = ROOT / "openkhuzdulgraph_amd" / "khuzdul" / "_examples"


@pytest.mark.skipif(not EX_DIR.is_dir(), reason="example apps were not built")
@pytest.mark.parametrize("method", ["auto", "plain"])
def test_bounded_set_op_examples_print_exact_counts(method):
    """intersect_and_count, bounded_intersect_with and bounded_subtracted_by
    inside apply() (src/vertex_set.cc:88-184), through this package's
    example apps: printed counts equal the reference binaries' triangle and
    4-clique counts and the open-wedge closed form on the golden fixtures."""
    from test_khuzdul_bounded_ops import FIXTURES, expected

    for name in FIXTURES:
        want = expected(name)
        for app in sorted(want):
            print(f"[case] app={app} graph={name} method={method}", flush=True)
            env = dict(os.environ)
            if method == "plain":
                env["KZ_METHOD"] = "plain"
            out = subprocess.run([str(EX_DIR / app), str(GOLDEN / f"{name}.dgraph")], env=env,
                                 capture_output=True, text=True, timeout=120)
            assert out.returncode == 0, out.stderr[-2000:]
            got = {int(m.group(1)): int(m.group(2)) for m in PAT.finditer(out.stdout)}
            assert [got[i] for i in sorted(got)] == want[app][0], (app, name, out.stdout[-1500:])
