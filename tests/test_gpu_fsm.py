"""GPU parity of labelled frequent subgraph mining (kz_fsm_mine) against
the oracle's restatement of kautomine/three_fsm.cc: every candidate's MNI
support, per-level frequent counts and support sums, on the golden labelled
fixtures (whose per-level results the reference binary printed), on
medium graphs with skewed degrees and few labels, and on labels set on a
graph generated on the GPU."""

import json
from pathlib import Path

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

GOLD_DIR = Path(__file__).parent / "golden"
FSM_GOLD = json.loads((GOLD_DIR / "golden_fsm.json").read_text())


def _check(dg, rowptr, col, labels, threshold, tag):
    print(f"[case] fsm {tag} threshold={threshold}", flush=True)
    got = dg.fsm(threshold)
    want = oracle.fsm(rowptr, col, labels, threshold)
    assert got["num_candidates"] == want["num_candidates"], tag
    assert got["candidates"] == want["candidates"], tag
    assert got["counts"] == want["counts"] and got["sums"] == want["sums"], tag
    return got


@pytest.mark.parametrize("name", sorted(FSM_GOLD))
def test_fsm_golden_fixtures(kzdev, name):
    kz = kzdev
    e = FSM_GOLD[name]
    g = kz.load_dgraph(str(GOLD_DIR / e["file"]))
    dg = g.to_device()  # labels travel with the .dgraph
    for t, ref in e["three_fsm"].items():
        got = _check(dg, g.rowptr, g.col, g.labels, int(t), name)
        assert got["counts"] == ref["counts"] and got["sums"] == ref["sums"], (name, t)
    dg.release()


# (name, graph builder, labels, thresholds); sizes keep the oracle's literal
# enumeration to seconds
MEDIUM = [
    ("chunglu_8k_l5", lambda kz: kz.gen_chung_lu(8000, 40000, 2.3, 800, 9), 5, (30, 200)),
    ("rmat11_ef8_l3", lambda kz: kz.gen_rmat(11, 8, seed=4), 3, (60, 300)),
    ("er_3k_l8", lambda kz: kz.gen_erdos_renyi(3000, 30000, 3), 8, (20,)),
    ("dense_l1", lambda kz: kz.gen_erdos_renyi(200, 8000, 8), 1, (10,)),
]


@pytest.mark.parametrize("case", range(len(MEDIUM)))
def test_fsm_medium_graphs(kzdev, case):
    kz = kzdev
    name, make, nl, thresholds = MEDIUM[case]
    g = make(kz)
    lab = np.random.default_rng(case).integers(0, nl, g.num_vertices).astype(np.uint32)
    dg = g.to_device()
    dg.set_labels(lab)
    for t in thresholds:
        _check(dg, g.rowptr, g.col, lab, t, name)
    # relabelling replaces the labelled lists
    lab2 = (lab + 1) % (int(lab.max()) + 1)
    dg.set_labels(lab2)
    _check(dg, g.rowptr, g.col, lab2, thresholds[0], name + "_relabelled")
    dg.release()


def test_fsm_on_gpu_generated_graph(kzdev):
    kz = kzdev
    dg = kz.gpu_gen_chung_lu(8000, 40000, 2.3, 800, 9)
    host = dg.download()
    lab = np.random.default_rng(4).integers(0, 4, dg.num_vertices).astype(np.uint32)
    dg.set_labels(lab)
    _check(dg, host.rowptr, host.col, lab, 30, "gpu_chunglu_8k_l4")
    dg.release()


def test_fsm_errors(kzdev):
    kz = kzdev
    g = kz.gen_erdos_renyi(100, 400, 2)
    dg = g.to_device()
    with pytest.raises(kz.KzError):
        dg.fsm(5)  # no labels yet
    with pytest.raises(kz.KzError):
        dg.set_labels(np.full(g.num_vertices, 48, np.uint32))  # label out of range
    o = g.oriented().to_device()
    with pytest.raises(kz.KzError):
        o.set_labels(np.zeros(g.num_vertices, np.uint32))  # symmetric graphs only
    dg.release()
    o.release()
