"""ctypes binding of include/kz.h (libkz_hip.so).

The library is loaded from openkhuzdulgraph_amd/lib/ only; there is no
fallback implementation: a missing library raises ImportError-like
KzLibraryMissing at first use.
"""

from __future__ import annotations

import ctypes as C
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "lib" / "libkz_hip.so"

KZ_OK = 0
KZ_ERR_INVALID = -1
KZ_ERR_NOMEM = -2
KZ_ERR_IO = -3
KZ_ERR_DEVICE = -4
KZ_ERR_UNSUPPORTED = -5
KZ_ERR_BUDGET = -6
KZ_ERR_STATE = -7

KZ_MAX_PATTERN = 5
KZ_MAX_AGGREGATORS = 8
KZ_MAX_OPS = 4
KZ_MAX_OUTS = 4
KZ_MAX_STEP_LISTS = 128
KZ_CENSUS_PATTERNS = 8
KZ_QUERY_CLIQUE = 1
KZ_QUERY_CENSUS = 2
KZ_PENDING = 1
CENSUS_PATTERNS = ("wedge", "triangle", "3-path", "3-star", "4-cycle", "tailed-triangle", "diamond", "4-clique")

KZ_METHOD_PLAIN = 0
KZ_METHOD_PERSISTENT = 1

KZ_SET_NONE = 0
KZ_SET_NBR = 1
KZ_SET_SLOT = 2
KZ_OP_INTERSECT = 1
KZ_OP_SUBTRACT = 2
KZ_OUT_SCATTER = 1
KZ_OUT_COUNT = 2


class KzLibraryMissing(RuntimeError):
    """libkz_hip.so has not been built (python -m openkhuzdulgraph_amd.build)."""


class KzError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"kz error {code}: {msg}")
        self.code = code


class kz_csr(C.Structure):
    _fields_ = [
        ("num_vertices", C.c_uint32),
        ("max_degree", C.c_uint32),
        ("num_entries", C.c_uint64),
        ("oriented", C.c_uint32),
        ("has_labels", C.c_uint32),
        ("rowptr", C.POINTER(C.c_uint64)),
        ("col", C.POINTER(C.c_uint32)),
        ("labels", C.POINTER(C.c_uint32)),
    ]


class kz_device_info(C.Structure):
    _fields_ = [
        ("name", C.c_char * 64),
        ("arch", C.c_char * 32),
        ("compute_units", C.c_int),
        ("warp_size", C.c_int),
        ("total_mem", C.c_uint64),
        ("lds_per_block", C.c_uint64),
        ("clock_khz", C.c_int),
        ("mem_clock_khz", C.c_int),
        ("mem_bus_width", C.c_int),
        ("l2_bytes", C.c_int),
    ]


class kz_run_stats(C.Structure):
    _fields_ = [
        ("device_ms", C.c_double),
        ("wall_ms", C.c_double),
        ("launches", C.c_uint64),
        ("tasks", C.c_uint64),
        ("embeddings", C.c_uint64 * (KZ_MAX_PATTERN + 1)),
        ("chunks", C.c_uint64),
        ("peak_bytes", C.c_uint64),
    ]

    def as_dict(self) -> dict:
        return {
            "device_ms": self.device_ms,
            "wall_ms": self.wall_ms,
            "launches": int(self.launches),
            "tasks": int(self.tasks),
            "embeddings": [int(x) for x in self.embeddings],
            "chunks": int(self.chunks),
            "peak_bytes": int(self.peak_bytes),
        }


class kz_bucket_cfg(C.Structure):
    _fields_ = [("bound", C.c_uint32 * 4), ("block", C.c_uint32 * 5), ("serial", C.c_uint32)]

    def as_dict(self) -> dict:
        return {"bound": list(self.bound), "block": list(self.block), "serial": int(self.serial)}


class kz_paged_cfg(C.Structure):
    _fields_ = [("list_bytes", C.c_uint64), ("cache_bytes", C.c_uint64), ("cache_degree_threshold", C.c_uint32),
                ("reserved", C.c_uint32)]


class kz_paged_stats(C.Structure):
    _fields_ = [("device_ms", C.c_double), ("wall_ms", C.c_double), ("batches", C.c_uint64),
                ("requests", C.c_uint64), ("request_refs", C.c_uint64), ("host_bytes", C.c_uint64),
                ("cache_bytes", C.c_uint64), ("cached_vertices", C.c_uint64)]

    def as_dict(self) -> dict:
        return {name: (float(getattr(self, name)) if name.endswith("_ms") else int(getattr(self, name)))
                for name, _ in self._fields_}


KZ_FSM_LEVELS = 6
FSM_LEVELS = ("vertex", "edge", "wedge", "triangle", "four_chain", "three_star")


class kz_fsm_stats(C.Structure):
    _fields_ = [
        ("num_frequent", C.c_uint32 * KZ_FSM_LEVELS),
        ("support_sum", C.c_uint64 * KZ_FSM_LEVELS),
        ("num_candidates", C.c_uint32 * KZ_FSM_LEVELS),
        ("num_labels", C.c_uint32),
        ("device_ms", C.c_double),
        ("total_ms", C.c_double),
    ]


class kz_fsm_candidate(C.Structure):
    _fields_ = [("level", C.c_uint32), ("labels", C.c_uint32 * 4), ("support", C.c_uint32),
                ("frequent", C.c_uint32)]


class kz_server_cfg(C.Structure):
    _fields_ = [("hbm_budget", C.c_uint64), ("max_inflight", C.c_uint32), ("reserved", C.c_uint32)]


class kz_query_result(C.Structure):
    _fields_ = [
        ("counts", C.c_uint64 * 8),
        ("queue_ms", C.c_double),
        ("run_ms", C.c_double),
        ("total_ms", C.c_double),
        ("status", C.c_int),
        ("reserved", C.c_int),
    ]

    def as_dict(self) -> dict:
        return {"counts": [int(x) for x in self.counts], "queue_ms": self.queue_ms, "run_ms": self.run_ms,
                "total_ms": self.total_ms, "status": int(self.status)}


class kz_set_ref(C.Structure):
    _fields_ = [("kind", C.c_uint8), ("a", C.c_uint8), ("b", C.c_uint8), ("reserved", C.c_uint8)]


class kz_op(C.Structure):
    _fields_ = [("op", C.c_uint8), ("reserved", C.c_uint8 * 3), ("lhs", kz_set_ref), ("rhs", kz_set_ref)]


class kz_out(C.Structure):
    _fields_ = [
        ("kind", C.c_uint8),
        ("aggregator", C.c_uint8),
        ("child_active", C.c_uint8),
        ("finalized", C.c_uint8),
        ("src", kz_set_ref),
        ("pred", C.c_uint32),
    ]


class kz_step_list(C.Structure):
    _fields_ = [
        ("num_ops", C.c_uint8),
        ("num_outs", C.c_uint8),
        ("reserved", C.c_uint8 * 2),
        ("ops", kz_op * KZ_MAX_OPS),
        ("outs", kz_out * KZ_MAX_OUTS),
    ]


class kz_plan_desc(C.Structure):
    _fields_ = [
        ("pattern_size", C.c_uint32),
        ("num_aggregators", C.c_uint32),
        ("multiplicity", C.c_uint32 * KZ_MAX_AGGREGATORS),
        ("num_step_lists", C.c_uint32),
        ("lists", kz_step_list * KZ_MAX_STEP_LISTS),
        ("type_table", (C.c_uint8 * 256) * KZ_MAX_PATTERN),
    ]


_lib = None

_SIGS = {
    "kz_last_error": (C.c_char_p, []),
    "kz_dgraph_load": (C.c_int, [C.c_char_p, C.POINTER(kz_csr)]),
    "kz_dgraph_store": (C.c_int, [C.c_char_p, C.POINTER(kz_csr)]),
    "kz_csr_release": (None, [C.POINTER(kz_csr)]),
    "kz_csr_from_edges": (C.c_int, [C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.c_uint64, C.c_uint32,
                                    C.POINTER(kz_csr)]),
    "kz_gen_erdos_renyi": (C.c_int, [C.c_uint32, C.c_uint64, C.c_uint64, C.POINTER(kz_csr)]),
    "kz_gen_rmat": (C.c_int, [C.c_uint32, C.c_uint32, C.c_double, C.c_double, C.c_double, C.c_uint64,
                              C.POINTER(kz_csr)]),
    "kz_gen_chung_lu": (C.c_int, [C.c_uint32, C.c_uint64, C.c_double, C.c_uint32, C.c_uint64, C.POINTER(kz_csr)]),
    "kz_orient_by_degree": (C.c_int, [C.POINTER(kz_csr), C.POINTER(kz_csr)]),
    "kz_csr_check": (C.c_int, [C.POINTER(kz_csr)]),
    "kz_csr_digest": (C.c_uint64, [C.POINTER(kz_csr)]),
    "kz_csr_set_labels": (C.c_int, [C.POINTER(kz_csr), C.POINTER(C.c_uint32)]),
    "kz_init": (C.c_int, [C.c_int]),
    "kz_shutdown": (C.c_int, []),
    "kz_get_device_info": (C.c_int, [C.POINTER(kz_device_info)]),
    "kz_graph_upload": (C.c_int, [C.POINTER(kz_csr), C.POINTER(C.c_void_p)]),
    "kz_graph_release": (C.c_int, [C.c_void_p]),
    "kz_graph_num_vertices": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint32)]),
    "kz_gpu_gen_erdos_renyi": (C.c_int, [C.c_uint32, C.c_uint64, C.c_uint64, C.c_int, C.POINTER(C.c_void_p)]),
    "kz_gpu_gen_rmat": (C.c_int, [C.c_uint32, C.c_uint32, C.c_double, C.c_double, C.c_double, C.c_uint64, C.c_int,
                                  C.POINTER(C.c_void_p)]),
    "kz_gpu_gen_chung_lu": (C.c_int, [C.c_uint32, C.c_uint64, C.c_double, C.c_uint32, C.c_uint64, C.c_int,
                                      C.POINTER(C.c_void_p)]),
    "kz_graph_download": (C.c_int, [C.c_void_p, C.POINTER(kz_csr)]),
    "kz_graph_orient": (C.c_int, [C.c_void_p, C.POINTER(C.c_void_p)]),
    "kz_paged_create": (C.c_int, [C.POINTER(kz_csr), C.POINTER(kz_paged_cfg), C.POINTER(C.c_void_p)]),
    "kz_paged_count_cliques": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_uint64), C.POINTER(kz_paged_stats)]),
    "kz_paged_release": (C.c_int, [C.c_void_p]),
    "kz_graph_info": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint32), C.POINTER(C.c_uint64), C.POINTER(C.c_uint32),
                                C.POINTER(C.c_int)]),
    "kz_set_sort_chunk": (C.c_int, [C.c_uint64]),
    "kz_count_cliques": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.POINTER(C.c_uint64), C.POINTER(kz_run_stats)]),
    "kz_plan_clique": (C.c_int, [C.c_int, C.POINTER(kz_plan_desc)]),
    "kz_plan_validate": (C.c_int, [C.POINTER(kz_plan_desc)]),
    "kz_count_plan": (C.c_int, [C.c_void_p, C.POINTER(kz_plan_desc), C.POINTER(C.c_uint64), C.c_uint32,
                                C.POINTER(kz_run_stats)]),
    "kz_set_level_budget": (C.c_int, [C.c_uint64]),
    "kz_clique_workload": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint64)]),
    "kz_motif_census": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint64), C.POINTER(kz_run_stats)]),
    "kz_graph_set_labels": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint32)]),
    "kz_fsm_domains": (C.c_int, [C.c_void_p, C.c_uint32, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]),
    "kz_fsm_mine": (C.c_int, [C.c_void_p, C.c_uint32, C.POINTER(kz_fsm_stats), C.POINTER(kz_fsm_candidate),
                              C.c_uint32, C.POINTER(C.c_uint32)]),
    "kz_server_start": (C.c_int, [C.POINTER(kz_server_cfg), C.POINTER(C.c_void_p)]),
    "kz_server_submit": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.POINTER(C.c_uint64)]),
    "kz_server_poll": (C.c_int, [C.c_void_p, C.c_uint64, C.c_int, C.POINTER(kz_query_result)]),
    "kz_server_stop": (C.c_int, [C.c_void_p]),
    "kz_clique_get_buckets": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(kz_bucket_cfg)]),
    "kz_clique_set_buckets": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(kz_bucket_cfg)]),
    "kz_clique_tune": (C.c_int, [C.c_void_p, C.c_int, C.c_uint32, C.POINTER(kz_bucket_cfg), C.POINTER(C.c_double)]),
    "kz_batch_create": (C.c_int, [C.POINTER(C.c_void_p), C.c_uint32, C.POINTER(C.c_int), C.c_uint32,
                                  C.POINTER(C.c_void_p)]),
    "kz_batch_run": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint64), C.POINTER(kz_run_stats)]),
    "kz_batch_tune": (C.c_int, [C.c_void_p, C.c_uint32, C.POINTER(kz_bucket_cfg), C.POINTER(C.c_double)]),
    "kz_batch_release": (C.c_int, [C.c_void_p]),
    "kz_set_op_batch": (C.c_int, [C.c_int, C.POINTER(C.c_uint32), C.POINTER(C.c_uint64), C.c_uint32,
                                  C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.c_uint32,
                                  C.POINTER(C.c_uint32), C.POINTER(C.c_uint64), C.POINTER(C.c_uint32)]),
    "kz_set_op_batch_noalloc": (C.c_int, [C.c_int, C.POINTER(C.c_uint32), C.POINTER(C.c_uint64), C.c_uint32,
                                  C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.c_uint32,
                                  C.POINTER(C.c_uint32), C.POINTER(C.c_uint64), C.POINTER(C.c_uint32)]),
}

KZ_SETOP_INTERSECT = 1
KZ_SETOP_SUBTRACT = 2
KZ_SETOP_INTERSECT_COUNT = 3
KZ_SETOP_TRIM_LOWER_BOUND = 4
KZ_SETOP_NOALLOC_MAX_BYTES = 32 << 20


def lib() -> C.CDLL:
    """Load libkz_hip.so (once) and bind every declared symbol."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise KzLibraryMissing(f"{LIB_PATH} not built; run python -m openkhuzdulgraph_amd.build")
        L = C.CDLL(str(LIB_PATH))
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def exported_symbols() -> list[str]:
    return list(_SIGS)


def check(rc: int) -> None:
    if rc != KZ_OK:
        msg = lib().kz_last_error()
        raise KzError(rc, msg.decode() if msg else "")
