"""Drop-in boundary for labelled mining, CPU side: the reference's
kautomine/three_fsm.cc, compiled unchanged against the Khuzdul headers,
is probed run by run (KZ_PROBE_ONLY=1; every pattern then reports the
largest support, so every level's candidates are probed).  Each run's
apply() (three_fsm.cc:52-250) must come out as the candidate's labelled
shape: the edge, wedge (center first), triangle, four-chain v2 - v0 - v1
- v3 and three-star (center v1) of the reference's generators
(:276-604), with the labels the generators produce."""

import itertools
import os
import subprocess
from pathlib import Path

import pytest

import openkhuzdulgraph_amd as kz

ROOT = Path(__file__).resolve().parent.parent
APP = ROOT / "openkhuzdulgraph_amd" / "khuzdul" / "_apps" / "kautomine_three_fsm"
GOLDEN = ROOT / "tests" / "golden"

# kz_fsm_mine levels and the adjacency of the app's pattern vertices
SHAPES = {1: [2, 1], 2: [6, 1, 1], 3: [6, 5, 3], 4: [6, 9, 1, 2], 5: [2, 13, 2, 2]}


def expected_candidates(L: int):
    """Every candidate the generators build when every pattern is frequent
    (three_fsm.cc:276-604 with all subpatterns present)."""
    r = range(L)
    out = []
    out += [(1, (a, b)) for a in r for b in r if a <= b]
    out += [(2, (a, b, c)) for a in r for b in r for c in r if b <= c]
    out += [(3, (a, b, c)) for a in r for b in r for c in r if a <= b <= c]
    out += [(4, (a, b, c, d)) for a in r for b in r for c in r for d in r if a <= b]
    out += [(5, (a, b, c, d)) for a, b, c, d in itertools.product(r, r, r, r)
            if b <= c <= d]  # generator order (l0, l1, l2, l3) = (center, leaf, leaf, leaf)
    return out


@pytest.mark.skipif(not APP.is_file(), reason="apps are built from the reference sources")
@pytest.mark.parametrize("fixture", ["dense_40_600_s4_l2.dgraph", "er_300_3000_s1_l4.dgraph"])
def test_three_fsm_runs_probe_to_their_candidate_shapes(fixture, tmp_path):
    g = kz.load_dgraph(str(GOLDEN / fixture))
    L = int(max(g.labels)) + 1
    trace = tmp_path / "trace.txt"
    env = dict(os.environ, KZ_PROBE_ONLY="1", KZ_LABELLED_TRACE=str(trace))
    out = subprocess.run([str(APP), str(GOLDEN / fixture), "1", "1"], env=env, capture_output=True, text=True,
                         timeout=300)
    assert out.returncode == 0, out.stderr[-2000:]
    got = []
    for line in trace.read_text().splitlines():
        f = line.split()
        level = int(f[1])
        k = len(SHAPES[level])
        labels = tuple(int(x) for x in f[3:3 + k])
        adj = [int(x) for x in f[4 + k:4 + 2 * k]]
        assert adj == SHAPES[level], line
        got.append((level, labels))
    want = expected_candidates(L)
    # the three-star generator writes (leaf, center, leaf, leaf) = (b, a, c, d)
    want = [(lv, (l[1], l[0], l[2], l[3]) if lv == 5 else l) for lv, l in want]
    assert sorted(got) == sorted(want)
