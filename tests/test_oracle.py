"""Pins the oracle (CPU restatement) before it is trusted as the checker.

* set primitives against Python set semantics and bisect on random and
  adversarial lists (src/vertex_set.cc:20-184, include/utilities.h:64-86);
* app schedules against the REFERENCE binaries' own counts on the committed
  fixtures (tests/golden/golden_counts.json, written by make_golden.py from
  the apps built from the reference sources);
* clique counts against brute-force subset enumeration and C(n, k) on K_n;
* 3-motif schedules against closed forms (kautomine: induced wedges =
  sum C(d,2) - 3T; kgraphpi pattern 1: sum C(d,2), non-induced);
* the merge-model byte pass against its definition on a tiny graph.
"""

import bisect
import json
import random
from math import comb
from pathlib import Path

import numpy as np
import pytest

import openkhuzdulgraph_amd as kz
import oracle

GOLDEN = Path(__file__).resolve().parent / "golden"


def sorted_unique(rng, n, hi):
    return sorted(rng.sample(range(hi), min(n, hi)))


def list_cases():
    rng = random.Random(1)
    top = 2**32 - 1
    cases = [([], []), ([], [1, 2]), ([1, 2], []), ([1, 2, 3], [1, 2, 3]), ([1, 3, 5], [2, 4, 6]),
             ([top - 2, top - 1], [top - 1]), ([0], [0]), (list(range(0, 200000, 2)), [3, 99998, 100000]),
             ([5], list(range(100000)))]
    for _ in range(200):
        a = sorted_unique(rng, rng.randint(0, 80), rng.choice([100, 1000, 2**32 - 1]))
        b = sorted_unique(rng, rng.randint(0, 80), rng.choice([100, 1000, 2**32 - 1]))
        cases.append((a, b))
    return cases


def test_set_ops_against_set_semantics():
    for a, b in list_cases():
        sa, sb = set(a), set(b)
        assert oracle.intersect(a, b).tolist() == sorted(sa & sb)
        assert oracle.subtract(a, b).tolist() == sorted(sa - sb)
        assert oracle.intersect_count(a, b) == len(sa & sb)
        for ub in (0, 1, 50, 500, 2**32 - 1):
            # the bounded variants stop at the first element of a >= ub
            cut = next((i for i, x in enumerate(a) if x >= ub), len(a))
            assert oracle.bounded_intersect(a, b, ub).tolist() == [x for x in a[:cut] if x in sb]
            assert oracle.bounded_subtract(a, b, ub).tolist() == [x for x in a[:cut] if x not in sb]


def test_trim_against_bisect():
    rng = random.Random(2)
    for _ in range(300):
        a = sorted_unique(rng, rng.randint(0, 50), 200)
        x = rng.randint(0, 210)
        assert oracle.trim(a, x) == bisect.bisect_left(a, x)
        assert oracle.trim_lower_bound(a, x) == bisect.bisect_right(a, x)


def _golden():
    return json.loads((GOLDEN / "golden_counts.json").read_text())


@pytest.mark.parametrize("name", list(_golden()))
def test_apps_match_reference_binaries(name):
    e = _golden()[name]
    g = kz.load_dgraph(str(GOLDEN / e["file"]))
    for app, want in e["counts"].items():
        if not want:  # the reference aborted on this input (multiplicity assert)
            continue
        got = oracle.run_app(app, g.rowptr, g.col)[1]
        assert got == want, (name, app, got, want)


def test_cliques_brute_force_and_closed_forms():
    for seed in range(4):
        g = kz.gen_erdos_renyi(48, 500, seed)
        for k, app in ((3, "triangle"), (4, "four_clique"), (5, "five_clique")):
            bf = oracle.brute_cliques(g.rowptr, g.col, k)
            assert bf == oracle.py_cliques(g.rowptr, g.col, k)
            for fam in ("kautomine", "kgraphpi"):
                assert oracle.run_app(f"{fam}_{app}", g.rowptr, g.col)[1] == [bf]
            o = g.oriented()
            assert oracle.run_app(f"kautomine_{app}", o.rowptr, o.col)[1] == [bf]
    # complete graph K_12: C(12, k)
    n = 12
    src, dst = zip(*[(i, j) for i in range(n) for j in range(i + 1, n)])
    K = kz.from_edges(np.array(src), np.array(dst), n)
    for k, app in ((3, "triangle"), (4, "four_clique"), (5, "five_clique")):
        assert oracle.run_app(f"kautomine_{app}", K.rowptr, K.col)[1] == [comb(n, k)]


def test_three_motif_closed_forms():
    for seed in range(3):
        g = kz.gen_chung_lu(300, 1500, 2.3, 40, seed)
        d = np.diff(g.rowptr).astype(np.int64)
        wedges_noninduced = int((d * (d - 1) // 2).sum())
        T = oracle.brute_cliques(g.rowptr, g.col, 3) if g.num_vertices <= 256 else \
            oracle.run_app("kautomine_triangle", g.rowptr, g.col)[1][0]
        ka = oracle.run_app("kautomine_three_motif", g.rowptr, g.col)[1]
        kg = oracle.run_app("kgraphpi_three_motif", g.rowptr, g.col)[1]
        assert ka == [T, wedges_noninduced - 3 * T]
        assert kg == [T, wedges_noninduced]
    small = kz.gen_erdos_renyi(40, 200, 5)
    bm = oracle.brute_motifs(small.rowptr, small.col)
    assert oracle.run_app("kautomine_three_motif", small.rowptr, small.col)[1] == [bm[1], bm[0]]


def test_clique_bytes_definition():
    g = kz.gen_erdos_renyi(30, 150, 7).oriented()
    rp, col = g.rowptr.astype(np.int64), g.col.astype(np.int64)
    n = len(rp) - 1
    N = [set(col[rp[v]:rp[v + 1]].tolist()) for v in range(n)]
    m = int(rp[-1])
    s3 = sum(len(N[u]) + len(N[v]) for u in range(n) for v in N[u])
    T = s4 = 0
    for u in range(n):
        for v in N[u]:
            c = N[u] & N[v]
            T += len(c)
            s4 += len(c) ** 2 + sum(len(N[w]) for w in c)
    b = oracle.clique_bytes(g.rowptr, g.col, kmax=4)
    assert b["B3"] == 4 * (m + s3)
    assert b["B4"] == 4 * (m + s3) + 4 * (T + s4)
    assert b["triangles"] == T
    # root subsets add up to the whole
    half = np.arange(0, n, 2, dtype=np.uint32)
    other = np.arange(1, n, 2, dtype=np.uint32)
    assert oracle.clique_bytes(g.rowptr, g.col, 4, roots=half)["B4"] + \
        oracle.clique_bytes(g.rowptr, g.col, 4, roots=other)["B4"] == b["B4"]


def test_motif_brute_force_identities():
    """Induced 4-vertex census against the non-induced counts through the
    standard inclusion matrix (3-stars = sum C(d,3), degree_analysis.cc:46)."""
    g = kz.gen_erdos_renyi(30, 160, 11)
    w, t, path, star, cyc, paw, dia, k4 = oracle.brute_motifs(g.rowptr, g.col)
    d = np.diff(g.rowptr).astype(np.int64)
    assert star + paw + 2 * dia + 4 * k4 == int((d * (d - 1) * (d - 2) // 6).sum())
    assert k4 == oracle.brute_cliques(g.rowptr, g.col, 4)
    assert t == oracle.brute_cliques(g.rowptr, g.col, 3)
    assert w + 3 * t == int((d * (d - 1) // 2).sum())


def _sym(n, edges):
    src = [a for a, b in edges] + [b for a, b in edges]
    dst = [b for a, b in edges] + [a for a, b in edges]
    return kz.from_edges(np.array(src, np.uint32), np.array(dst, np.uint32), n)


def test_census_against_brute_force():
    """The combinatorial census (any size) against exhaustive enumeration of
    every 3- and 4-subset on random graphs from empty to complete."""
    rng = np.random.default_rng(3)
    for _ in range(120):
        n = int(rng.integers(4, 48))
        m = int(rng.integers(1, n * (n - 1) // 2 + 1))
        s, t = rng.integers(0, n, m), rng.integers(0, n, m)
        g = kz.from_edges(np.concatenate([s, t]), np.concatenate([t, s]), n)
        if g.num_vertices == 0:
            continue
        assert oracle.motif_census(g.rowptr, g.col) == oracle.brute_motifs(g.rowptr, g.col)


def test_census_closed_forms():
    """Each pattern alone, K_n, stars and cycles: induced counts by hand."""
    order = ["wedge", "triangle", "3-path", "3-star", "4-cycle", "tailed", "diamond", "4-clique"]

    def census(n, edges):
        g = _sym(n, edges)
        return dict(zip(order, oracle.motif_census(g.rowptr, g.col)))

    only = {
        "3-path": (4, [(0, 1), (1, 2), (2, 3)]),
        "3-star": (4, [(0, 1), (0, 2), (0, 3)]),
        "4-cycle": (4, [(0, 1), (1, 2), (2, 3), (3, 0)]),
        "tailed": (4, [(0, 1), (1, 2), (2, 0), (2, 3)]),
        "diamond": (4, [(0, 1), (1, 2), (2, 0), (1, 3), (2, 3)]),
        "4-clique": (4, [(a, b) for a in range(4) for b in range(a + 1, 4)]),
    }
    for name, (n, edges) in only.items():
        c = census(n, edges)
        assert all(c[k] == (1 if k == name else 0) for k in order[2:]), (name, c)
    for n in (5, 9, 17):
        c = census(n, [(a, b) for a in range(n) for b in range(a + 1, n)])
        assert c["4-clique"] == comb(n, 4) and c["triangle"] == comb(n, 3)
        assert sum(c[k] for k in order) == comb(n, 4) + comb(n, 3)
        c = census(n, [(0, b) for b in range(1, n)])
        assert c["3-star"] == comb(n - 1, 3) and c["wedge"] == comb(n - 1, 2)
        c = census(n, [(a, (a + 1) % n) for a in range(n)])
        assert c["3-path"] == n and c["wedge"] == n and c["4-cycle"] == 0


def test_census_non_induced_identities():
    """Non-induced counts against independent formulas on a Chung-Lu graph:
    stars from degrees, 4-cliques from the clique oracle, and the induced
    counts re-expanded through the containment matrix."""
    g = kz.gen_chung_lu(3000, 15000, 2.3, 200, 5)
    ind, non = oracle.motif_census(g.rowptr, g.col, non_induced=True)
    d = np.diff(g.rowptr).astype(np.int64)
    o = g.oriented()
    assert non[3] == int((d * (d - 1) * (d - 2) // 6).sum())
    assert non[0] == int((d * (d - 1) // 2).sum())
    assert non[7] == oracle.run_app("kautomine_four_clique", o.rowptr, o.col)[1][0]
    assert non[1] == oracle.run_app("kautomine_triangle", o.rowptr, o.col)[1][0]
    w, t, p3, st, c4, paw, dia, k4 = ind
    A = [[1, 0, 4, 2, 6, 12], [0, 1, 0, 1, 2, 4], [0, 0, 1, 0, 1, 3], [0, 0, 0, 1, 4, 12], [0, 0, 0, 0, 1, 6],
         [0, 0, 0, 0, 0, 1]]
    x = [p3, st, c4, paw, dia, k4]
    assert [sum(A[i][j] * x[j] for j in range(6)) for i in range(6)] == non[2:]


# --------------------------------------------------------------- three_fsm --
FSM_GOLD = json.loads((Path(__file__).parent / "golden" / "golden_fsm.json").read_text())
FSM_TEMPLATES = {  # level -> (pattern size, edges between positions)
    1: (2, [(0, 1)]),
    2: (3, [(0, 1), (0, 2)]),
    3: (3, [(0, 1), (0, 2), (1, 2)]),
    4: (4, [(0, 1), (0, 2), (1, 3)]),
    5: (4, [(1, 0), (1, 2), (1, 3)]),
}


@pytest.mark.parametrize("name", sorted(FSM_GOLD))
def test_oracle_fsm_matches_reference_golden(name):
    """Frequent-pattern counts and support sums per level equal what the
    reference kautomine three_fsm binary printed for the same labelled
    .dgraph and threshold."""
    e = FSM_GOLD[name]
    g = kz.load_dgraph(str(Path(__file__).parent / "golden" / e["file"]))
    assert g.labels is not None and int(g.labels.max()) + 1 == e["num_labels"]
    for t, want in e["three_fsm"].items():
        got = oracle.fsm(g.rowptr, g.col, g.labels, int(t))
        assert got["counts"] == want["counts"], (name, t)
        assert got["sums"] == want["sums"], (name, t)


def _brute_support(n, adj, lab, level, labels):
    """MNI support of one labelled template by brute force: every injective
    assignment of graph vertices to pattern positions with the template's
    edges present and the labels matching."""
    from itertools import permutations
    size, edges = FSM_TEMPLATES[level]
    dom = [set() for _ in range(size)]
    for t in permutations(range(n), size):
        if all(lab[t[i]] == labels[i] for i in range(size)) and all(t[b] in adj[t[a]] for a, b in edges):
            for i in range(size):
                dom[i].add(t[i])
    return min([n] + [len(d) for d in dom])


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_oracle_fsm_brute_force(seed):
    """Every candidate's support equals a brute-force minimum-image count
    over injective embeddings on small random labelled graphs (thresholds
    low enough that every level has candidates)."""
    rng = np.random.default_rng(seed)
    g = kz.gen_erdos_renyi(11, 30, seed)
    n = g.num_vertices
    lab = rng.integers(0, 2 + seed % 2, n).astype(np.uint32)
    rp, col = g.rowptr, g.col
    adj = [set(int(x) for x in col[rp[v]:rp[v + 1]]) for v in range(n)]
    r = oracle.fsm(rp, col, lab, 1)
    assert r["num_candidates"] > 10
    for level, labels, sup, fr in r["candidates"]:
        assert sup == _brute_support(n, adj, lab, level, labels), (level, labels)
        assert fr == (sup >= 1)
