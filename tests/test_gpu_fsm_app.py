"""Drop-in boundary for labelled mining on the GPU: the reference's
kautomine/three_fsm.cc, compiled unchanged against libkhuzdul_amd.so,
mines the golden labelled fixtures and prints, per level, the number of
frequent patterns and their support sum that the REFERENCE binary printed
(tests/golden/golden_fsm.json, made by tests/golden/make_golden.py --fsm).
Each of the app's run() calls is probed into its labelled pattern and the
pattern's domains come from kz_fsm_domains."""

import json
import os
import re
import subprocess
from pathlib import Path

import pytest

pytestmark = pytest.mark.gpu

ROOT = Path(__file__).resolve().parent.parent
APP = ROOT / "openkhuzdulgraph_amd" / "khuzdul" / "_apps" / "kautomine_three_fsm"
GOLDEN = ROOT / "tests" / "golden"
FSM_GOLD = json.loads((GOLDEN / "golden_fsm.json").read_text())
LEVELS = ["single-vertex", "single-edge", "wedge", "triangle", "four-chain", "three-star"]
LINE = re.compile(r"Number of frequent ([a-z-]+) patterns: (\d+), support sum(?: of them)?: (\d+)")


@pytest.mark.skipif(not APP.is_file(), reason="apps are built from the reference sources")
@pytest.mark.parametrize("name", sorted(FSM_GOLD))
def test_unchanged_three_fsm_prints_reference_results(name):
    e = FSM_GOLD[name]
    for t, ref in sorted(e["three_fsm"].items()):
        print(f"[case] three_fsm {name} threshold={t}", flush=True)
        out = subprocess.run([str(APP), str(GOLDEN / e["file"]), str(t), "1"], env=dict(os.environ),
                             capture_output=True, text=True, timeout=300)
        assert out.returncode == 0, out.stderr[-2000:]
        got = {m.group(1): (int(m.group(2)), int(m.group(3))) for m in LINE.finditer(out.stdout)}
        assert [got.get(lv, (None, None))[0] for lv in LEVELS] == ref["counts"], (name, t, out.stdout[-2000:])
        assert [got.get(lv, (None, None))[1] for lv in LEVELS] == ref["sums"], (name, t, out.stdout[-2000:])
