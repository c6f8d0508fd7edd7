"""Host graph layer: generators, clean-up, orientation, .dgraph I/O, checks.

Rules checked against independent numpy restatements of
datasets/live_journal/standardize.cc:66-114 (clean-up),
toolkits/converter.cc:237-263 (both directions, sorted) and
toolkits/to_oriented_dataset_converter.cc:27-124 (orientation).
"""

import ctypes as C
import json
import os
import struct
import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest

import openkhuzdulgraph_amd as kz
from openkhuzdulgraph_amd import _ffi

GOLDEN = Path(__file__).resolve().parent / "golden"


def np_clean(src, dst, n):
    """Independent clean-up: drop loops/dupes/isolated, relabel ascending."""
    s = np.asarray(src, np.int64)
    d = np.asarray(dst, np.int64)
    keep = s != d
    s, d = s[keep], d[keep]
    a, b = np.minimum(s, d), np.maximum(s, d)
    pairs = np.unique(a * n + b)
    a, b = pairs // n, pairs % n
    used = np.zeros(n, bool)
    used[a] = True
    used[b] = True
    newid = np.cumsum(used) - 1
    a, b = newid[a], newid[b]
    n2 = int(used.sum())
    both_s = np.concatenate([a, b])
    both_d = np.concatenate([b, a])
    order = np.lexsort((both_d, both_s))
    both_s, both_d = both_s[order], both_d[order]
    rowptr = np.zeros(n2 + 1, np.uint64)
    np.add.at(rowptr, both_s + 1, 1)
    return np.cumsum(rowptr).astype(np.uint64), both_d.astype(np.uint32)


def np_orient(rowptr, col):
    n = len(rowptr) - 1
    deg = np.diff(rowptr).astype(np.int64)
    order = np.lexsort((np.arange(n), deg))  # (deg asc, id asc)
    rank = np.empty(n, np.int64)
    rank[order] = np.arange(n)
    src = np.repeat(np.arange(n), deg)
    rs, rd = rank[src], rank[col.astype(np.int64)]
    keep = rd > rs
    rs, rd = rs[keep], rd[keep]
    o = np.lexsort((rd, rs))
    rs, rd = rs[o], rd[o]
    rp = np.zeros(n + 1, np.uint64)
    np.add.at(rp, rs + 1, 1)
    return np.cumsum(rp).astype(np.uint64), rd.astype(np.uint32)


def test_from_edges_cleanup_rules():
    rng = np.random.default_rng(0)
    n = 500
    src = rng.integers(0, n, 4000)
    dst = rng.integers(0, n, 4000)
    src[:50] = dst[:50]  # self loops
    src = np.concatenate([src, dst[:100]])  # reversed duplicates
    dst = np.concatenate([dst, src[:100]])
    g = kz.from_edges(src, dst, n)
    rp, col = np_clean(src, dst, n)
    assert np.array_equal(g.rowptr, rp)
    assert np.array_equal(g.col, col)
    g.check()


def test_orientation_rules():
    g = kz.gen_rmat(10, 16, seed=4)
    o = g.oriented()
    rp, col = np_orient(g.rowptr, g.col)
    assert np.array_equal(o.rowptr, rp)
    assert np.array_equal(o.col, col)
    assert o.oriented_flag and o.num_entries * 2 == g.num_entries
    o.check()


def test_generators_deterministic_across_thread_counts():
    code = ("import sys; sys.path.insert(0, %r); import openkhuzdulgraph_amd as kz; "
            "print(kz.gen_rmat(12, 8, seed=9).digest(), kz.gen_erdos_renyi(5000, 40000, 9).digest(), "
            "kz.gen_chung_lu(5000, 20000, 2.3, 300, 9).digest())") % str(Path(__file__).resolve().parent.parent)
    outs = set()
    for t in ("1", "3", "8"):
        env = dict(os.environ, OMP_NUM_THREADS=t)
        outs.add(subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True,
                                check=True).stdout.strip())
    assert len(outs) == 1


def test_generator_shapes():
    er = kz.gen_erdos_renyi(1 << 12, 1 << 15, 1)
    er.check()
    assert er.num_vertices <= 1 << 12
    assert abs(er.num_entries / 2 - (1 << 15)) < 0.02 * (1 << 15)  # few duplicates at this density
    r = kz.gen_rmat(12, 16, seed=1)
    r.check()
    assert r.max_degree > 20 * r.num_entries / r.num_vertices  # skewed
    cl = kz.gen_chung_lu(20000, 60000, 2.3, 500, 1)
    cl.check()
    assert cl.max_degree <= 2 * 500


def test_dgraph_roundtrip_and_layout(tmp_path):
    g = kz.gen_erdos_renyi(200, 1000, 3)
    p = tmp_path / "g.dgraph"
    g.store(str(p))
    raw = p.read_bytes()
    nv, ne, hl, he = struct.unpack("<IQii", raw[:20])  # graph_loader.h:37-42
    assert (nv, ne, hl, he) == (g.num_vertices, g.num_entries, 0, 0)
    edges = np.frombuffer(raw[20:], dtype=np.uint32).reshape(-1, 2)
    assert len(edges) == g.num_entries
    assert np.array_equal(edges[:, 1], g.col)
    assert np.all(np.diff(edges[:, 0].astype(np.int64)) >= 0)
    h = kz.load_dgraph(str(p))
    assert h.digest() == g.digest() and not h.oriented_flag
    o = g.oriented()
    o.store(str(p))
    ho = kz.load_dgraph(str(p))
    assert ho.oriented_flag and ho.digest() == o.digest()


def test_dgraph_rejects_bad_files(tmp_path):
    p = tmp_path / "bad.dgraph"
    p.write_bytes(struct.pack("<IQii", 4, 2, 0, 0) + struct.pack("<4I", 1, 2, 0, 1))  # unsorted
    with pytest.raises(kz.KzError):
        kz.load_dgraph(str(p))
    p.write_bytes(struct.pack("<IQii", 4, 3, 0, 0) + struct.pack("<4I", 0, 1, 1, 0))  # size mismatch
    with pytest.raises(kz.KzError):
        kz.load_dgraph(str(p))
    p.write_bytes(struct.pack("<IQii", 4, 1, 0, 1) + struct.pack("<2I", 0, 1))  # edge labels
    with pytest.raises(kz.KzError):
        kz.load_dgraph(str(p))
    with pytest.raises(kz.KzError):
        kz.load_dgraph(str(tmp_path / "missing.dgraph"))


def _csr_from_arrays(rowptr, col, oriented=0):
    rp = np.ascontiguousarray(rowptr, np.uint64)
    cl = np.ascontiguousarray(col, np.uint32)
    s = _ffi.kz_csr()
    s.num_vertices = len(rp) - 1
    s.num_entries = len(cl)
    s.oriented = oriented
    s.rowptr = rp.ctypes.data_as(C.POINTER(C.c_uint64))
    s.col = cl.ctypes.data_as(C.POINTER(C.c_uint32))
    return s, (rp, cl)


@pytest.mark.parametrize("rowptr,col,oriented", [
    ([0, 1, 1], [0], 0),                 # self loop
    ([0, 2, 3, 4], [2, 1, 0, 0], 0),     # unsorted list
    ([0, 1, 1], [1], 0),                 # missing reverse edge
    ([0, 0, 1], [0], 1),                 # oriented edge pointing down
    ([0, 1, 2], [5, 0], 0),              # id out of range
])
def test_csr_check_catches(rowptr, col, oriented):
    s, keep = _csr_from_arrays(rowptr, col, oriented)
    assert _ffi.lib().kz_csr_check(C.byref(s)) == _ffi.KZ_ERR_INVALID


def test_golden_fixture_digests():
    golden = json.loads((GOLDEN / "golden_counts.json").read_text())
    for name, e in golden.items():
        g = kz.load_dgraph(str(GOLDEN / e["file"]))
        assert f"{g.digest():016x}" == e["digest"], name
        assert g.num_vertices == e["num_vertices"] and g.num_entries == e["num_entries"]
        g.check()
