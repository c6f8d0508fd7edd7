"""Oracle: CPU restatement of the OpenKhuzdulGraph embedding-extension path.

TEST INFRASTRUCTURE ONLY.  Imported by tests/, by __graft_entry__.smoke() as
the checker, and by bench.py's cpu_baseline leg.  The product package
(openkhuzdulgraph_amd) never imports this module.

The C++ restatement (kz_oracle.cpp, built by oracle/Makefile into
libkz_oracle.so) follows src/vertex_set.cc:20-184, include/utilities.h:64-86
and every app's generated apply() (kautomine/*.cc, kgraphpi/*.cc); see the
citations in kz_oracle.cpp.  The pure-Python helpers below (set operations,
brute force, closed forms) pin the C++ restatement in tests/test_oracle.py.
"""

from __future__ import annotations

import ctypes as C
import itertools
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "libkz_oracle.so"
REF_DIR = HERE / "_ref"

APPS = {
    "kautomine_triangle": 0,
    "kautomine_four_clique": 1,
    "kautomine_five_clique": 2,
    "kautomine_three_motif": 3,
    "kgraphpi_triangle": 4,
    "kgraphpi_four_clique": 5,
    "kgraphpi_five_clique": 6,
    "kgraphpi_three_motif": 7,
}

_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        L = C.CDLL(str(LIB_PATH))
        U32P = C.POINTER(C.c_uint32)
        U64P = C.POINTER(C.c_uint64)
        for name in ("kzo_intersect", "kzo_subtract"):
            getattr(L, name).argtypes = [U32P, C.c_uint32, U32P, C.c_uint32, U32P]
            getattr(L, name).restype = C.c_uint32
        L.kzo_intersect_count.argtypes = [U32P, C.c_uint32, U32P, C.c_uint32]
        L.kzo_intersect_count.restype = C.c_uint32
        for name in ("kzo_bounded_intersect", "kzo_bounded_subtract"):
            getattr(L, name).argtypes = [U32P, C.c_uint32, U32P, C.c_uint32, C.c_uint32, U32P]
            getattr(L, name).restype = C.c_uint32
        for name in ("kzo_trim", "kzo_trim_lower_bound"):
            getattr(L, name).argtypes = [U32P, C.c_uint32, C.c_uint32]
            getattr(L, name).restype = C.c_uint32
        L.kzo_run_app.argtypes = [C.c_int, U64P, U32P, C.c_uint32, U32P, C.c_uint64, C.c_int, U64P, U64P]
        L.kzo_run_app.restype = C.c_int
        L.kzo_brute_cliques.argtypes = [U64P, U32P, C.c_uint32, C.c_int]
        L.kzo_brute_cliques.restype = C.c_longlong
        L.kzo_brute_motifs.argtypes = [U64P, U32P, C.c_uint32, U64P]
        L.kzo_brute_motifs.restype = C.c_int
        L.kzo_clique_bytes.argtypes = [U64P, U32P, C.c_uint32, C.c_int, C.c_int, U32P, C.c_uint64, U64P]
        L.kzo_clique_bytes.restype = C.c_int
        L.kzo_motif_census.argtypes = [U64P, U32P, C.c_uint32, C.c_int, U64P, U64P]
        L.kzo_fsm.argtypes = [U64P, U32P, C.c_uint32, U32P, C.c_uint32, U32P, U64P, U32P, C.c_uint32, U32P]
        L.kzo_motif_census.restype = C.c_int
        _lib = L
    return _lib


def _u32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.uint32)


def _p32(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_uint32))


def _p64(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_uint64))


# --------------------------------------------------------------- set ops --

def intersect(a, b) -> np.ndarray:
    a, b = _u32(a), _u32(b)
    out = np.zeros(max(1, min(len(a), len(b))), np.uint32)
    k = lib().kzo_intersect(_p32(a), len(a), _p32(b), len(b), _p32(out))
    return out[:k]


def subtract(a, b) -> np.ndarray:
    a, b = _u32(a), _u32(b)
    out = np.zeros(max(1, len(a)), np.uint32)
    k = lib().kzo_subtract(_p32(a), len(a), _p32(b), len(b), _p32(out))
    return out[:k]


def intersect_count(a, b) -> int:
    a, b = _u32(a), _u32(b)
    return int(lib().kzo_intersect_count(_p32(a), len(a), _p32(b), len(b)))


def bounded_intersect(a, b, ub: int) -> np.ndarray:
    a, b = _u32(a), _u32(b)
    out = np.zeros(max(1, min(len(a), len(b))), np.uint32)
    k = lib().kzo_bounded_intersect(_p32(a), len(a), _p32(b), len(b), ub, _p32(out))
    return out[:k]


def bounded_subtract(a, b, ub: int) -> np.ndarray:
    a, b = _u32(a), _u32(b)
    out = np.zeros(max(1, len(a)), np.uint32)
    k = lib().kzo_bounded_subtract(_p32(a), len(a), _p32(b), len(b), ub, _p32(out))
    return out[:k]


def trim(a, ub: int) -> int:
    a = _u32(a)
    return int(lib().kzo_trim(_p32(a), len(a), ub))


def trim_lower_bound(a, lb: int) -> int:
    a = _u32(a)
    return int(lib().kzo_trim_lower_bound(_p32(a), len(a), lb))


# ------------------------------------------------------------ app counts --

def run_app(app: str, rowptr, col, roots=None, threads: int = 0):
    """(raw aggregator sums, printed counts) of one app over the roots."""
    rp = np.ascontiguousarray(rowptr, dtype=np.uint64)
    cl = _u32(col) if len(col) else np.zeros(1, np.uint32)
    n = len(rp) - 1
    raw = np.zeros(2, np.uint64)
    out = np.zeros(2, np.uint64)
    if roots is None:
        rptr, nroots = None, n
    else:
        r = _u32(roots)
        rptr, nroots = _p32(r), len(r)
    k = lib().kzo_run_app(APPS[app], _p64(rp), _p32(cl), n, rptr, nroots, threads, _p64(raw), _p64(out))
    if k < 0:
        raise ValueError(app)
    return [int(x) for x in raw[:k]], [int(x) for x in out[:k]]


def brute_cliques(rowptr, col, k: int) -> int:
    rp = np.ascontiguousarray(rowptr, dtype=np.uint64)
    cl = _u32(col) if len(col) else np.zeros(1, np.uint32)
    return int(lib().kzo_brute_cliques(_p64(rp), _p32(cl), len(rp) - 1, k))


def brute_motifs(rowptr, col) -> list[int]:
    """[wedge, triangle, 3-path, 3-star, 4-cycle, tailed-triangle, diamond, 4-clique] (induced)."""
    rp = np.ascontiguousarray(rowptr, dtype=np.uint64)
    cl = _u32(col) if len(col) else np.zeros(1, np.uint32)
    out = np.zeros(8, np.uint64)
    if lib().kzo_brute_motifs(_p64(rp), _p32(cl), len(rp) - 1, _p64(out)) < 0:
        raise ValueError("graph too large for brute force")
    return [int(x) for x in out]


def motif_census(rowptr, col, threads: int = 0, non_induced: bool = False):
    """Induced census [wedge, triangle, 3-path, 3-star, 4-cycle, tailed-triangle,
    diamond, 4-clique] of a symmetric graph (combinatorial, any size); with
    non_induced=True also returns the non-induced counts in the same order."""
    rp = np.ascontiguousarray(rowptr, dtype=np.uint64)
    cl = _u32(col) if len(col) else np.zeros(1, np.uint32)
    out = np.zeros(8, np.uint64)
    ni = np.zeros(8, np.uint64)
    lib().kzo_motif_census(_p64(rp), _p32(cl), len(rp) - 1, threads, _p64(out), _p64(ni))
    res = [int(x) for x in out]
    return (res, [int(x) for x in ni]) if non_induced else res


def clique_bytes(rowptr, col, kmax: int = 4, threads: int = 0, roots=None) -> dict:
    """Merge-model bytes B3/B4/B5 (and clique counts) of the roots given
    (all vertices when None) on an oriented graph."""
    rp = np.ascontiguousarray(rowptr, dtype=np.uint64)
    cl = _u32(col) if len(col) else np.zeros(1, np.uint32)
    out = np.zeros(6, np.uint64)
    if roots is None:
        rptr, nr = None, 0
    else:
        r = _u32(roots)
        rptr, nr = _p32(r), len(r)
    lib().kzo_clique_bytes(_p64(rp), _p32(cl), len(rp) - 1, kmax, threads, rptr, nr, _p64(out))
    return {"B3": int(out[0]), "B4": int(out[1]), "B5": int(out[2]),
            "triangles": int(out[3]), "k4": int(out[4]), "k5": int(out[5])}


# ------------------------------------------- pure-Python pins (small only) --

def py_cliques(rowptr, col, k: int) -> int:
    """k-cliques by itertools over an adjacency set (tiny graphs only)."""
    n = len(rowptr) - 1
    adj = [set(int(x) for x in col[rowptr[v]:rowptr[v + 1]]) for v in range(n)]
    for v in range(n):
        for u in list(adj[v]):
            adj[u].add(v)
    return sum(1 for S in itertools.combinations(range(n), k)
               if all(b in adj[a] for a, b in itertools.combinations(S, 2)))


def reference_binary(app: str) -> Path:
    return REF_DIR / app


def run_plan(plan, rowptr, col, threads: int = 0) -> list[int]:
    """Raw aggregator sums of a kz_plan_desc (ctypes struct or bytes)."""
    import ctypes as _C

    rp = np.ascontiguousarray(rowptr, dtype=np.uint64)
    cl = _u32(col) if len(col) else np.zeros(1, np.uint32)
    raw = np.zeros(8, np.uint64)
    buf = bytes(plan) if not isinstance(plan, (bytes, bytearray)) else bytes(plan)
    cbuf = _C.create_string_buffer(buf, len(buf))
    L = lib()
    L.kzo_run_plan.argtypes = [_C.c_void_p, _C.POINTER(_C.c_uint64), _C.POINTER(_C.c_uint32), _C.c_uint32,
                               _C.c_int, _C.POINTER(_C.c_uint64)]
    L.kzo_run_plan.restype = _C.c_int
    L.kzo_run_plan(_C.cast(cbuf, _C.c_void_p), _p64(rp), _p32(cl), len(rp) - 1, threads, _p64(raw))
    n = int.from_bytes(buf[4:8], "little")
    return [int(x) for x in raw[:n]]


FSM_LEVELS = ("vertex", "edge", "wedge", "triangle", "four_chain", "three_star")


def fsm(rowptr, col, labels, threshold: int, cap: int = 1 << 20) -> dict:
    """kautomine/three_fsm.cc restated: per level (FSM_LEVELS) the number of
    frequent patterns and their support sum, plus every evaluated candidate
    as (level, labels, support, frequent)."""
    L = lib()
    rp = np.ascontiguousarray(rowptr, dtype=np.uint64)
    cl = _u32(col)
    lab = _u32(labels)
    n = len(rp) - 1
    counts = np.zeros(6, np.uint32)
    sums = np.zeros(6, np.uint64)
    cand = np.zeros((cap, 7), np.uint32)
    nc = np.zeros(1, np.uint32)
    L.kzo_fsm(_p64(rp), _p32(cl), n, _p32(lab), threshold, _p32(counts), _p64(sums), _p32(cand), cap, _p32(nc))
    k = min(int(nc[0]), cap)
    cands = [(int(r[0]), tuple(int(x) for x in r[1:5]), int(r[5]), bool(r[6])) for r in cand[:k]]
    return {"counts": [int(x) for x in counts], "sums": [int(x) for x in sums], "candidates": cands,
            "num_candidates": int(nc[0])}

