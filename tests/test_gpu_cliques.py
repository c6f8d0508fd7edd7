"""GPU parity of k-clique counting (persistent and plain forms) against the
oracle's restatement of the kautomine/kgraphpi schedules."""

import pytest

import oracle

pytestmark = pytest.mark.gpu

APP = {3: "kautomine_triangle", 4: "kautomine_four_clique", 5: "kautomine_five_clique"}
KG = {3: "kgraphpi_triangle", 4: "kgraphpi_four_clique", 5: "kgraphpi_five_clique"}


def _graphs(kz):
    yield "er_sym", kz.gen_erdos_renyi(2000, 40000, 5)
    yield "er_orient", kz.gen_erdos_renyi(2000, 40000, 5).oriented()
    yield "rmat12_orient", kz.gen_rmat(12, 16, seed=7).oriented()
    yield "rmat12_sym", kz.gen_rmat(12, 16, seed=7)
    yield "dense_orient", kz.gen_erdos_renyi(300, 30000, 9).oriented()


@pytest.mark.parametrize("k", [3, 4, 5])
def test_cliques_match_oracle(kzdev, k):
    kz = kzdev
    for name, g in _graphs(kz):
        print(f"[case] graph={name} k={k} n={g.num_vertices} entries={g.num_entries}", flush=True)
        want = oracle.run_app(APP[k], g.rowptr, g.col)[1][0]
        assert oracle.run_app(KG[k], g.rowptr, g.col)[1][0] == want
        dg = g.to_device()
        got_p = dg.count_cliques(k, "persistent")
        got_l = dg.count_cliques(k, "plain")
        assert got_p == want, (name, k, got_p, want)
        assert got_l == want, (name, k, got_l, want)
        # replays of the captured graph must agree with the first run
        for _ in range(3):
            assert dg.count_cliques(k, "persistent") == want
        dg.release()


def test_plain_chunking_is_exact(kzdev):
    kz = kzdev
    g = kz.gen_rmat(12, 16, seed=3).oriented()
    want = oracle.run_app(APP[4], g.rowptr, g.col)[1][0]
    dg = g.to_device()
    kz.set_level_budget(1 << 20)
    try:
        got = dg.count_cliques(4, "plain")
        assert dg.last_stats["chunks"] > 1
    finally:
        kz.set_level_budget(4 << 30)
    assert got == want


def test_workload_terms_match_oracle(kzdev):
    kz = kzdev
    for g in (kz.gen_rmat(11, 16, seed=5).oriented(), kz.gen_erdos_renyi(1500, 30000, 2).oriented()):
        want = oracle.clique_bytes(g.rowptr, g.col, kmax=4)
        dg = g.to_device()
        got = dg.clique_workload()
        dg.release()
        assert got["B3"] == want["B3"]
        assert got["B4"] == want["B4"]
        assert got["triangles"] == want["triangles"]


def test_batch_matches_single_runs(kzdev):
    """Many (graph, k) queries in one hipGraph launch: each count equals the
    oracle's, and equals the graph's own single-query run."""
    kz = kzdev
    import random

    rng = random.Random(5)
    hosts = []
    for i in range(12):
        n = rng.choice([300, 1000, 3000])
        if i % 2:
            g = kz.gen_erdos_renyi(n, n * rng.choice([4, 16, 40]), 100 + i)
        else:
            g = kz.gen_rmat(rng.choice([9, 10, 11]), rng.choice([4, 16, 32]), seed=100 + i)
        hosts.append(g.oriented() if i % 3 else g)
    devs = [g.to_device() for g in hosts]
    b = kz.Batch(devs, [3, 4, 5])
    for rep in range(3):
        got = b.run()
        for gi, g in enumerate(hosts):
            for ki, (k, app) in enumerate(zip((3, 4, 5), (APP[3], APP[4], APP[5]))):
                want = oracle.run_app(app, g.rowptr, g.col)[1][0]
                assert int(got[gi, ki]) == want, (rep, gi, k)
    for gi, d in enumerate(devs):
        assert d.count_cliques(4, "persistent") == int(got[gi, 1])
    b.release()


def test_autotuner_keeps_counts_exact(kzdev):
    kz = kzdev
    g = kz.gen_rmat(13, 16, seed=2).oriented()
    want = oracle.run_app(APP[4], g.rowptr, g.col)[1][0]
    dg = g.to_device()
    res = dg.tune_buckets(4, reps=1)
    assert res["ms"] > 0 and res["cfg"]["bound"][0] in (32, 64)
    assert dg.count_cliques(4, "persistent") == want
    dg.set_buckets(4, [16, 128, 512, 1024], [256, 512, 512, 512, 512])
    assert dg.count_cliques(4, "persistent") == want
    with pytest.raises(kz.KzError):
        dg.set_buckets(4, [128, 256, 512, 1024], [256, 256, 256, 512, 512])
    dg.release()


def _oriented_by_id(kz, g):
    """The forward-by-id part of a symmetric graph (u -> v for v > u) as an
    oriented graph, through a .dgraph file (packed 20-byte header, then the
    sorted {src, dst} pairs): low-id hubs keep their forward degrees."""
    import tempfile

    import numpy as np
    rp, col = g.rowptr, g.col
    src = np.repeat(np.arange(g.num_vertices, dtype=np.uint32), np.diff(rp).astype(np.int64))
    keep = col > src
    pairs = np.empty((int(keep.sum()), 2), np.uint32)
    pairs[:, 0], pairs[:, 1] = src[keep], col[keep]
    header = np.zeros(1, dtype=[("nv", "<u4"), ("ne", "<u8"), ("nl", "<i4"), ("el", "<i4")])
    header["nv"], header["ne"] = g.num_vertices, len(pairs)
    with tempfile.NamedTemporaryFile(suffix=".dgraph") as f:
        f.write(header.tobytes())
        f.write(pairs.tobytes())
        f.flush()
        o = kz.load_dgraph(f.name)
    assert o.oriented_flag
    return o


@pytest.mark.parametrize("k", [3, 4, 5])
def test_hub_and_bucket_paths_exact(kzdev, k):
    """Every bucket path (wave, LDS rows, hub rows in HBM + per-row count) on
    the same graphs: default bounds, bounds that push most roots into the hub
    kernels, and a dense graph whose forward degrees exceed 1024."""
    kz = kzdev
    cases = [("rmat12", kz.gen_rmat(12, 16, seed=21).oriented()),
             ("er_dense", kz.gen_erdos_renyi(600, 40000, 22).oriented())]
    # two planted hubs with forward degree > 1024 on a sparse background
    import numpy as np
    rng = np.random.default_rng(23)
    src = np.concatenate([rng.integers(0, 3000, 60000), np.zeros(2000, np.int64), np.ones(1500, np.int64)])
    dst = np.concatenate([rng.integers(0, 3000, 60000), np.arange(1, 2001), np.arange(2, 1502)])
    cases.append(("planted_hubs", _oriented_by_id(kz, kz.from_edges(src, dst, 3000))))
    # two planted roots with forward degree in (512, 1024]: the last LDS
    # bucket, whose 1024-thread workgroups keep packed rows and probe queues
    src = np.concatenate([rng.integers(0, 3000, 60000), np.zeros(900, np.int64), np.ones(700, np.int64)])
    dst = np.concatenate([rng.integers(0, 3000, 60000), np.arange(1, 901), np.arange(2, 702)])
    cases.append(("planted_mid", _oriented_by_id(kz, kz.from_edges(src, dst, 3000))))
    # a planted hub of forward degree 5,000: its hash and list leave one
    # 1024-thread rows workgroup per CU, which then keeps probe queues beside
    # an 8-bit-per-id filter
    src = np.concatenate([rng.integers(0, 6000, 60000), np.zeros(5000, np.int64)])
    dst = np.concatenate([rng.integers(0, 6000, 60000), np.arange(1, 5001)])
    cases.append(("planted_big_hub", _oriented_by_id(kz, kz.from_edges(src, dst, 6000))))
    for name, g in cases:
        print(f"[case] graph={name} k={k} max_degree={g.max_degree}", flush=True)
        want = oracle.run_app(APP[k], g.rowptr, g.col)[1][0]
        dg = g.to_device()
        assert dg.count_cliques(k, "persistent") == want, (name, "default")
        dg.set_buckets(k, [16, 32, 48, 64], [256, 256, 256, 512, 512])
        assert dg.count_cliques(k, "persistent") == want, (name, "hub-heavy")
        dg.set_buckets(k, [8, 24, 40, 56], [256, 512, 512, 512, 256])
        assert dg.count_cliques(k, "persistent") == want, (name, "hub-heavy-256")
        # both launch policies and 1024-thread workgroups on every bucket
        for serial in (False, True):
            dg.set_buckets(k, [64, 256, 512, 1024], [256, 1024, 1024, 1024, 1024], serial)
            assert dg.count_cliques(k, "persistent") == want, (name, "1024", serial)
            dg.set_buckets(k, [32, 96, 160, 224], [256, 256, 512, 1024, 512], serial)
            assert dg.count_cliques(k, "persistent") == want, (name, "mixed", serial)
        dg.release()


_NO_QUEUE_CHILD = r"""
import sys
sys.path.insert(0, sys.argv[1])
import numpy as np
import openkhuzdulgraph_amd as kz
sys.path.insert(0, sys.argv[1] + "/tests")
from test_gpu_cliques import _oriented_by_id
kz.init(0)
rng = np.random.default_rng(23)
src = np.concatenate([rng.integers(0, 3000, 60000), np.zeros(900, np.int64), np.ones(700, np.int64)])
dst = np.concatenate([rng.integers(0, 3000, 60000), np.arange(1, 901), np.arange(2, 702)])
graphs = {"planted_mid": _oriented_by_id(kz, kz.from_edges(src, dst, 3000)),
          "rmat12": kz.gen_rmat(12, 16, seed=21).oriented()}
for name, g in graphs.items():
    dg = g.to_device()
    print(name, *[dg.count_cliques(k, "persistent") for k in (3, 4, 5)], flush=True)
    dg.release()
"""


def test_kernels_without_probe_queues_are_exact(tmp_path):
    """KZ_STREAM_QUEUE=0 (read once per process, so in a child process):
    the bucket kernels stream without probe queues and with full rows,
    instances the default configuration no longer launches."""
    import os
    import subprocess
    import sys
    from pathlib import Path

    import numpy as np

    import openkhuzdulgraph_amd as kz

    root = Path(__file__).resolve().parent.parent
    print("[case] KZ_STREAM_QUEUE=0 child: planted_mid, rmat12, k = 3..5", flush=True)
    out = subprocess.run([sys.executable, "-c", _NO_QUEUE_CHILD, str(root)], capture_output=True, text=True,
                         timeout=240, env=dict(os.environ, KZ_STREAM_QUEUE="0"))
    assert out.returncode == 0, out.stderr[-2000:]
    got = {line.split()[0]: [int(x) for x in line.split()[1:]] for line in out.stdout.splitlines() if line.strip()}
    rng = np.random.default_rng(23)
    src = np.concatenate([rng.integers(0, 3000, 60000), np.zeros(900, np.int64), np.ones(700, np.int64)])
    dst = np.concatenate([rng.integers(0, 3000, 60000), np.arange(1, 901), np.arange(2, 702)])
    graphs = {"planted_mid": _oriented_by_id(kz, kz.from_edges(src, dst, 3000)),
              "rmat12": kz.gen_rmat(12, 16, seed=21).oriented()}
    for name, g in graphs.items():
        want = [oracle.run_app(APP[k], g.rowptr, g.col)[1][0] for k in (3, 4, 5)]
        assert got[name] == want, (name, got[name], want)
