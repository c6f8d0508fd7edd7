"""openkhuzdulgraph_amd — MI355X-native embedding-extension engine.

Python host mirror of the C ABI in include/kz.h.  Everything here calls
libkz_hip.so; there is no Python or CPU compute path, and a missing library
raises KzLibraryMissing.

    import openkhuzdulgraph_amd as kz
    g = kz.gen_rmat(scale=18, edge_factor=16, seed=1).oriented()
    kz.init(0)
    dg = g.to_device()
    n4 = dg.count_cliques(4, method="persistent")
"""

from __future__ import annotations

import ctypes as C
from typing import Optional

import numpy as np

from . import _ffi
from ._ffi import KzError, KzLibraryMissing, check, lib

__all__ = [
    "Batch", "CSRGraph", "DeviceGraph", "KzError", "KzLibraryMissing", "init", "shutdown", "device_info",
    "gen_erdos_renyi", "gen_rmat", "gen_chung_lu", "from_edges", "load_dgraph", "set_level_budget",
    "clique_plan", "Server", "gpu_gen_erdos_renyi", "gpu_gen_rmat", "gpu_gen_chung_lu", "PagedGraph",
]

_METHODS = {"plain": _ffi.KZ_METHOD_PLAIN, "persistent": _ffi.KZ_METHOD_PERSISTENT}


class CSRGraph:
    """Host CSR owned by libkz_hip.so (kz_csr).  Arrays are exposed as
    zero-copy numpy views valid while this object lives."""

    def __init__(self, raw: _ffi.kz_csr):
        self._raw = raw

    def __del__(self):
        try:
            if self._raw.rowptr:
                lib().kz_csr_release(C.byref(self._raw))
        except Exception:
            pass

    @property
    def num_vertices(self) -> int:
        return int(self._raw.num_vertices)

    @property
    def num_entries(self) -> int:
        return int(self._raw.num_entries)

    @property
    def max_degree(self) -> int:
        return int(self._raw.max_degree)

    @property
    def oriented_flag(self) -> bool:
        return bool(self._raw.oriented)

    @property
    def rowptr(self) -> np.ndarray:
        return np.ctypeslib.as_array(self._raw.rowptr, (self.num_vertices + 1,))

    @property
    def col(self) -> np.ndarray:
        if self.num_entries == 0:
            return np.zeros(0, np.uint32)
        return np.ctypeslib.as_array(self._raw.col, (self.num_entries,))

    def degrees(self) -> np.ndarray:
        return np.diff(self.rowptr).astype(np.uint32)

    @property
    def labels(self):
        """Vertex labels (u32 view) or None."""
        if not self._raw.has_labels or not self._raw.labels:
            return None
        return np.ctypeslib.as_array(self._raw.labels, (max(1, self.num_vertices),))[: self.num_vertices]

    def set_labels(self, labels) -> "CSRGraph":
        """Attach vertex labels (copied; each below KZ_MAX_LABELS = 48)."""
        lab = np.ascontiguousarray(labels, dtype=np.uint32)
        if lab.shape != (self.num_vertices,):
            raise ValueError(f"need {self.num_vertices} labels, got {lab.shape}")
        check(lib().kz_csr_set_labels(C.byref(self._raw), lab.ctypes.data_as(C.POINTER(C.c_uint32))))
        return self

    def oriented(self) -> "CSRGraph":
        """Degree orientation (to_oriented_dataset_converter.cc:27-124)."""
        out = _ffi.kz_csr()
        check(lib().kz_orient_by_degree(C.byref(self._raw), C.byref(out)))
        return CSRGraph(out)

    def check(self) -> None:
        check(lib().kz_csr_check(C.byref(self._raw)))

    def digest(self) -> int:
        return int(lib().kz_csr_digest(C.byref(self._raw)))

    def store(self, path: str) -> None:
        check(lib().kz_dgraph_store(str(path).encode(), C.byref(self._raw)))

    def to_device(self) -> "DeviceGraph":
        h = C.c_void_p()
        check(lib().kz_graph_upload(C.byref(self._raw), C.byref(h)))
        return DeviceGraph(h, self)

    def __repr__(self) -> str:
        kind = "oriented" if self.oriented_flag else "symmetric"
        return f"CSRGraph(n={self.num_vertices}, entries={self.num_entries}, max_degree={self.max_degree}, {kind})"


def _new(fn, *args) -> CSRGraph:
    out = _ffi.kz_csr()
    check(fn(*args, C.byref(out)))
    return CSRGraph(out)


def gen_erdos_renyi(n: int, m: int, seed: int) -> CSRGraph:
    return _new(lib().kz_gen_erdos_renyi, n, m, seed)


def gen_rmat(scale: int, edge_factor: int = 16, a: float = 0.57, b: float = 0.19, c: float = 0.19,
             seed: int = 1) -> CSRGraph:
    return _new(lib().kz_gen_rmat, scale, edge_factor, a, b, c, seed)


def gen_chung_lu(n: int, num_draws: int, exponent: float, max_weight: int, seed: int) -> CSRGraph:
    return _new(lib().kz_gen_chung_lu, n, num_draws, exponent, max_weight, seed)


def from_edges(src, dst, num_vertices_hint: int) -> CSRGraph:
    s = np.ascontiguousarray(src, dtype=np.uint32)
    d = np.ascontiguousarray(dst, dtype=np.uint32)
    if s.shape != d.shape:
        raise ValueError("src and dst differ in length")
    P = C.POINTER(C.c_uint32)
    return _new(lib().kz_csr_from_edges, s.ctypes.data_as(P), d.ctypes.data_as(P), len(s), num_vertices_hint)


def load_dgraph(path: str) -> CSRGraph:
    return _new(lib().kz_dgraph_load, str(path).encode())


def _gpu_new(fn, *args) -> "DeviceGraph":
    h = C.c_void_p()
    check(fn(*args, C.byref(h)))
    return DeviceGraph(h)


def gpu_gen_erdos_renyi(n: int, m: int, seed: int, orient: bool = False) -> "DeviceGraph":
    """gen_erdos_renyi (and .oriented()) built on the GPU, resident in HBM."""
    return _gpu_new(lib().kz_gpu_gen_erdos_renyi, n, m, seed, int(orient))


def gpu_gen_rmat(scale: int, edge_factor: int = 16, a: float = 0.57, b: float = 0.19, c: float = 0.19,
                 seed: int = 1, orient: bool = False) -> "DeviceGraph":
    """gen_rmat (and .oriented()) built on the GPU, resident in HBM."""
    return _gpu_new(lib().kz_gpu_gen_rmat, scale, edge_factor, a, b, c, seed, int(orient))


def gpu_gen_chung_lu(n: int, num_draws: int, exponent: float, max_weight: int, seed: int,
                     orient: bool = False) -> "DeviceGraph":
    """gen_chung_lu (and .oriented()) built on the GPU, resident in HBM."""
    return _gpu_new(lib().kz_gpu_gen_chung_lu, n, num_draws, exponent, max_weight, seed, int(orient))


def set_sort_chunk(items: int) -> None:
    check(lib().kz_set_sort_chunk(int(items)))


def init(device: int = 0) -> None:
    check(lib().kz_init(device))


def shutdown() -> None:
    check(lib().kz_shutdown())


def device_info() -> dict:
    info = _ffi.kz_device_info()
    check(lib().kz_get_device_info(C.byref(info)))
    return {k: (getattr(info, k).decode() if isinstance(getattr(info, k), bytes) else getattr(info, k))
            for k, _ in info._fields_}


def set_level_budget(nbytes: int) -> None:
    check(lib().kz_set_level_budget(int(nbytes)))


def set_op_batch(op: str, lists, pairs, bounds=None, noalloc: bool = False):
    """Batched VertexSet operations on the GPU (kz_set_op_batch, or with
    noalloc=True kz_set_op_batch_noalloc: no device allocation, batches up
    to KZ_SETOP_NOALLOC_MAX_BYTES of staging).

    op: "intersect" | "subtract" | "intersect_count" | "trim_lower_bound".
    lists: sequence of strictly increasing id sequences; pairs: sequence of
    (a, b) list indices (b ignored by trim).  Returns a list of numpy arrays
    (intersect/subtract) or a numpy array of integers."""
    code = {"intersect": _ffi.KZ_SETOP_INTERSECT, "subtract": _ffi.KZ_SETOP_SUBTRACT,
            "intersect_count": _ffi.KZ_SETOP_INTERSECT_COUNT,
            "trim_lower_bound": _ffi.KZ_SETOP_TRIM_LOWER_BOUND}[op]
    arrs = [np.ascontiguousarray(x, dtype=np.uint32) for x in lists]
    off = np.zeros(len(arrs) + 1, np.uint64)
    off[1:] = np.cumsum([len(x) for x in arrs])
    vals = np.concatenate(arrs) if arrs and off[-1] else np.zeros(1, np.uint32)
    pa = np.ascontiguousarray([p[0] for p in pairs], dtype=np.uint32)
    pb = np.ascontiguousarray([p[1] for p in pairs], dtype=np.uint32)
    bd = np.ascontiguousarray(bounds if bounds is not None else np.zeros(len(pairs)), dtype=np.uint32)
    sizes = []
    for a, b in zip(pa, pb):
        na, nb = len(arrs[a]), len(arrs[b])
        sizes.append(min(na, nb) if code == _ffi.KZ_SETOP_INTERSECT else na)
    oo = np.zeros(len(pairs) + 1, np.uint64)
    oo[1:] = np.cumsum(sizes)
    outv = np.zeros(max(1, int(oo[-1])), np.uint32)
    outl = np.zeros(max(1, len(pairs)), np.uint32)
    P32, P64 = C.POINTER(C.c_uint32), C.POINTER(C.c_uint64)
    fn = lib().kz_set_op_batch_noalloc if noalloc else lib().kz_set_op_batch
    check(fn(code, vals.ctypes.data_as(P32), off.ctypes.data_as(P64), len(arrs),
             pa.ctypes.data_as(P32), pb.ctypes.data_as(P32), bd.ctypes.data_as(P32), len(pairs),
             outv.ctypes.data_as(P32), oo.ctypes.data_as(P64), outl.ctypes.data_as(P32)))
    if code in (_ffi.KZ_SETOP_INTERSECT, _ffi.KZ_SETOP_SUBTRACT):
        return [outv[int(oo[i]):int(oo[i]) + int(outl[i])].copy() for i in range(len(pairs))]
    return outl[:len(pairs)].copy()


class PagedGraph:
    """A degree-oriented graph whose neighbour lists stay in pinned host
    memory, counted in root batches with lists fetched into HBM on demand
    (kz_paged_*): list_bytes of HBM for one batch's lists, cache_bytes for
    the lists of the highest-degree vertices (total degree >=
    cache_degree_threshold)."""

    def __init__(self, graph: CSRGraph, list_bytes: int, cache_bytes: int = 0, cache_degree_threshold: int = 64):
        cfg = _ffi.kz_paged_cfg(int(list_bytes), int(cache_bytes), int(cache_degree_threshold), 0)
        self._h = C.c_void_p()
        check(lib().kz_paged_create(C.byref(graph._raw), C.byref(cfg), C.byref(self._h)))
        self.last_stats: dict = {}

    def count_cliques(self, k: int) -> int:
        out = C.c_uint64()
        st = _ffi.kz_paged_stats()
        check(lib().kz_paged_count_cliques(self._h, k, C.byref(out), C.byref(st)))
        self.last_stats = st.as_dict()
        return int(out.value)

    def release(self) -> None:
        if self._h:
            check(lib().kz_paged_release(self._h))
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass


class Batch:
    """Many (graph, k) clique queries in one hipGraph launch (kz_batch)."""

    def __init__(self, graphs, ks):
        self.graphs = list(graphs)  # keep the device graphs alive
        self.ks = list(ks)
        hs = (C.c_void_p * len(self.graphs))(*[g._h for g in self.graphs])
        karr = (C.c_int * len(self.ks))(*self.ks)
        self._h = C.c_void_p()
        check(lib().kz_batch_create(hs, len(self.graphs), karr, len(self.ks), C.byref(self._h)))
        self.last_stats: dict = {}

    def run(self) -> np.ndarray:
        out = (C.c_uint64 * (len(self.graphs) * len(self.ks)))()
        st = _ffi.kz_run_stats()
        check(lib().kz_batch_run(self._h, out, C.byref(st)))
        self.last_stats = st.as_dict()
        return np.array(out, dtype=np.uint64).reshape(len(self.graphs), len(self.ks))

    def tune(self, reps: int = 2) -> dict:
        cfg = _ffi.kz_bucket_cfg()
        ms = C.c_double()
        check(lib().kz_batch_tune(self._h, reps, C.byref(cfg), C.byref(ms)))
        return {"cfg": cfg.as_dict(), "ms": ms.value}

    def release(self) -> None:
        if self._h:
            check(lib().kz_batch_release(self._h))
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass


class Server:
    """Query server on resident graphs (kz_server_*): queries are admitted in
    arrival order within an HBM budget for prepared executions and launched
    concurrently on up to max_inflight streams."""

    def __init__(self, hbm_budget: int = 0, max_inflight: int = 4):
        cfg = _ffi.kz_server_cfg(hbm_budget, max_inflight, 0)
        self._h = C.c_void_p()
        check(lib().kz_server_start(C.byref(cfg), C.byref(self._h)))
        self._graphs = {}  # ticket -> graph (kept alive while the query runs)

    def submit(self, graph: "DeviceGraph", kind: str = "clique", k: int = 4) -> int:
        code = {"clique": _ffi.KZ_QUERY_CLIQUE, "census": _ffi.KZ_QUERY_CENSUS}[kind]
        t = C.c_uint64()
        check(lib().kz_server_submit(self._h, graph._h, code, k, C.byref(t)))
        self._graphs[int(t.value)] = graph
        return int(t.value)

    def poll(self, ticket: int, wait: bool = False):
        """Result dict when the query has finished, else None (wait=False)."""
        r = _ffi.kz_query_result()
        rc = lib().kz_server_poll(self._h, ticket, 1 if wait else 0, C.byref(r))
        if rc == _ffi.KZ_PENDING:
            return None
        check(rc)
        self._graphs.pop(ticket, None)
        res = r.as_dict()
        if res["status"]:
            res["error"] = lib().kz_last_error().decode(errors="replace")
        return res

    def stop(self) -> None:
        if self._h:
            check(lib().kz_server_stop(self._h))
            self._h = C.c_void_p()
            self._graphs.clear()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.stop()

    def __del__(self):
        try:
            self.stop()
        except Exception:
            pass


def clique_plan(k: int) -> _ffi.kz_plan_desc:
    p = _ffi.kz_plan_desc()
    check(lib().kz_plan_clique(k, C.byref(p)))
    return p


class DeviceGraph:
    """A CSR resident in HBM (kz_graph)."""

    def __init__(self, handle: C.c_void_p, host: Optional[CSRGraph] = None):
        self._h = handle
        self.host = host
        self.last_stats: dict = {}

    def release(self) -> None:
        if self._h:
            check(lib().kz_graph_release(self._h))
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass

    def _info(self):
        n, m, d, o = C.c_uint32(), C.c_uint64(), C.c_uint32(), C.c_int()
        check(lib().kz_graph_info(self._h, C.byref(n), C.byref(m), C.byref(d), C.byref(o)))
        return int(n.value), int(m.value), int(d.value), bool(o.value)

    @property
    def num_vertices(self) -> int:
        return self._info()[0]

    @property
    def num_entries(self) -> int:
        return self._info()[1]

    @property
    def max_degree(self) -> int:
        return self._info()[2]

    @property
    def oriented_flag(self) -> bool:
        return self._info()[3]

    def download(self) -> CSRGraph:
        """Copy of the resident graph as a host CSR."""
        out = _ffi.kz_csr()
        check(lib().kz_graph_download(self._h, C.byref(out)))
        return CSRGraph(out)

    def oriented(self) -> "DeviceGraph":
        """Degree orientation of this symmetric graph built on the GPU
        (kz_graph_orient): the same CSR as download().oriented()."""
        h = C.c_void_p()
        check(lib().kz_graph_orient(self._h, C.byref(h)))
        return DeviceGraph(h)

    def count_cliques(self, k: int, method: str = "persistent") -> int:
        out = C.c_uint64()
        st = _ffi.kz_run_stats()
        check(lib().kz_count_cliques(self._h, k, _METHODS[method], C.byref(out), C.byref(st)))
        self.last_stats = st.as_dict()
        return int(out.value)

    def tune_buckets(self, k: int, reps: int = 2) -> dict:
        """Autotune the degree-bucket thresholds of the persistent form."""
        cfg = _ffi.kz_bucket_cfg()
        ms = C.c_double()
        check(lib().kz_clique_tune(self._h, k, reps, C.byref(cfg), C.byref(ms)))
        return {"cfg": cfg.as_dict(), "ms": ms.value}

    def get_buckets(self, k: int) -> dict:
        cfg = _ffi.kz_bucket_cfg()
        check(lib().kz_clique_get_buckets(self._h, k, C.byref(cfg)))
        return cfg.as_dict()

    def set_buckets(self, k: int, bound, block, serial: bool = False) -> None:
        cfg = _ffi.kz_bucket_cfg()
        for i, x in enumerate(bound):
            cfg.bound[i] = x
        for i, x in enumerate(block):
            cfg.block[i] = x
        cfg.serial = 1 if serial else 0
        check(lib().kz_clique_set_buckets(self._h, k, C.byref(cfg)))

    def clique_workload(self) -> dict:
        """Merge-model byte terms (kz_clique_workload) and B3/B4 in bytes."""
        out = (C.c_uint64 * 6)()
        check(lib().kz_clique_workload(self._h, out))
        m, T, c2, sw, dsq, dv = (int(x) for x in out)
        b3 = 4 * (m + dsq + dv)
        return {"m": m, "triangles": T, "sum_c2": c2, "sum_tri_dw": sw, "sum_d2": dsq, "sum_dv": dv,
                "B3": b3, "B4": b3 + 4 * (T + c2 + sw)}

    def motif_census(self) -> list[int]:
        """Induced 3- and 4-vertex census (kz_motif_census) of a symmetric
        graph, in the order of _ffi.CENSUS_PATTERNS."""
        out = (C.c_uint64 * _ffi.KZ_CENSUS_PATTERNS)()
        st = _ffi.kz_run_stats()
        check(lib().kz_motif_census(self._h, out, C.byref(st)))
        self.last_stats = st.as_dict()
        return [int(x) for x in out]

    def set_labels(self, labels) -> "DeviceGraph":
        """Vertex labels for frequent subgraph mining (kz_graph_set_labels;
        symmetric graphs, each label below 48)."""
        lab = np.ascontiguousarray(labels, dtype=np.uint32)
        if lab.shape != (self.num_vertices,):
            raise ValueError(f"need {self.num_vertices} labels, got {lab.shape}")
        check(lib().kz_graph_set_labels(self._h, lab.ctypes.data_as(C.POINTER(C.c_uint32))))
        return self

    def fsm(self, threshold: int, cap: int = 1 << 20) -> dict:
        """Frequent labelled 1-4 vertex patterns with MNI support >= threshold
        (kz_fsm_mine, kautomine/three_fsm.cc): per level of _ffi.FSM_LEVELS
        the number of frequent patterns and their support sum (what the
        reference app prints), plus every evaluated candidate as (level,
        labels, support, frequent)."""
        st = _ffi.kz_fsm_stats()
        buf = (_ffi.kz_fsm_candidate * cap)()
        nc = C.c_uint32(0)
        check(lib().kz_fsm_mine(self._h, threshold, C.byref(st), buf, cap, C.byref(nc)))
        k = min(int(nc.value), cap)
        cands = [(int(c.level), tuple(int(x) for x in c.labels), int(c.support), bool(c.frequent)) for c in buf[:k]]
        self.last_stats = {"device_ms": st.device_ms, "total_ms": st.total_ms, "num_labels": int(st.num_labels),
                           "num_candidates": [int(x) for x in st.num_candidates]}
        return {"counts": [int(x) for x in st.num_frequent], "sums": [int(x) for x in st.support_sum],
                "candidates": cands, "num_candidates": int(nc.value)}

    def count_plan(self, plan: _ffi.kz_plan_desc) -> list[int]:
        n = max(1, int(plan.num_aggregators))
        out = (C.c_uint64 * _ffi.KZ_MAX_AGGREGATORS)()
        st = _ffi.kz_run_stats()
        check(lib().kz_count_plan(self._h, C.byref(plan), out, _ffi.KZ_MAX_AGGREGATORS, C.byref(st)))
        self.last_stats = st.as_dict()
        return [int(out[i]) for i in range(int(plan.num_aggregators))][:n]
