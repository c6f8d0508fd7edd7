"""Symmetric inputs of any degree (src/distributed_application.cc:322-326
enumerates every root; src/vertex_set.cc:20-49 intersects lists of any
length).  The clique kernels run on the degree orientation of a symmetric
graph, built on the GPU (kz_graph_orient): it must be the host's
orientation bit for bit, and the counts of a symmetric RMAT-16 -- whose
low-id hubs keep 9,794 higher-id neighbours -- must be the reference
binaries' counts through the C ABI, the unchanged apps, the batched form
and the server."""

import json
import os
import re
import subprocess
from pathlib import Path

import pytest

pytestmark = pytest.mark.gpu

ROOT = Path(__file__).resolve().parent.parent
APPS_DIR = ROOT / "openkhuzdulgraph_amd" / "khuzdul" / "_apps"
GOLDEN = json.loads((ROOT / "tests" / "golden" / "config_counts.json").read_text())["sym_rmat16_ef16"]
PAT = re.compile(r"Pattern count (\d+): (\d+)")


def _want(app):
    return GOLDEN["counts"][f"reference:{app}"]


@pytest.fixture(scope="module")
def sym16(kzdev):
    kz = kzdev
    from openkhuzdulgraph_amd import workloads as wl
    dg = wl.rmat(kz, 16, 16, gpu=True, orient=False)
    assert f"{dg.download().digest():016x}" == GOLDEN["digest"]
    yield dg
    dg.release()


@pytest.mark.parametrize("chunk", [1 << 30, 1 << 16])
def test_gpu_orientation_is_the_host_orientation(kzdev, chunk):
    kz = kzdev
    kz.set_sort_chunk(chunk)
    try:
        cases = [("er", kz.gen_erdos_renyi(1 << 14, 1 << 17, 3)),
                 ("rmat16", kz.gen_rmat(16, 16, seed=1)),
                 ("chunglu", kz.gen_chung_lu(30000, 200000, 2.3, 3000, 6)),
                 ("tiny", kz.gen_erdos_renyi(5, 3, 2))]
        for name, h in cases:
            print(f"[case] {name} chunk={chunk}", flush=True)
            dg = h.to_device()
            o = dg.oriented()
            got = o.download()
            want = h.oriented()
            assert got.oriented_flag
            assert (got.num_vertices, got.num_entries, got.max_degree) == \
                (want.num_vertices, want.num_entries, want.max_degree), name
            assert got.digest() == want.digest(), name
            o.release()
            dg.release()
    finally:
        kz.set_sort_chunk(1 << 30)


def test_symmetric_counts_through_the_abi(sym16):
    for method in ("persistent", "plain"):
        print(f"[case] sym rmat16 method={method}", flush=True)
        assert sym16.count_cliques(3, method) == _want("kautomine_triangle")
        assert sym16.count_cliques(4, method) == _want("kautomine_four_clique")


def test_symmetric_batch_and_server(kzdev, sym16):
    kz = kzdev
    print("[case] batch on symmetric rmat16", flush=True)
    b = kz.Batch([sym16], [3, 4])
    got = b.run()
    b.release()
    assert [int(x) for x in got[0]] == [_want("kautomine_triangle"), _want("kautomine_four_clique")]
    print("[case] server on symmetric rmat16", flush=True)
    with kz.Server(hbm_budget=8 << 30, max_inflight=2) as srv:
        t3, t4 = srv.submit(sym16, "clique", 3), srv.submit(sym16, "clique", 4)
        r3, r4 = srv.poll(t3, wait=True), srv.poll(t4, wait=True)
    assert r3["status"] == 0 and r3["counts"][0] == _want("kautomine_triangle")
    assert r4["status"] == 0 and r4["counts"][0] == _want("kautomine_four_clique")


@pytest.mark.skipif(not APPS_DIR.is_dir(), reason="app binaries were not built from the reference sources")
def test_unchanged_apps_on_symmetric_rmat16(sym16, tmp_path):
    path = tmp_path / "rmat16_sym.dgraph"
    sym16.download().store(str(path))
    for app in ("kautomine_triangle", "kgraphpi_triangle", "kautomine_four_clique", "kgraphpi_four_clique"):
        print(f"[case] app={app} symmetric rmat16", flush=True)
        out = subprocess.run([str(APPS_DIR / app), str(path), "1", "1"], env=dict(os.environ),
                             capture_output=True, text=True, timeout=120)
        assert out.returncode == 0, out.stderr[-2000:]
        got = {int(m.group(1)): int(m.group(2)) for m in PAT.finditer(out.stdout)}
        assert got.get(0) == _want(app), (app, out.stdout[-1500:])
