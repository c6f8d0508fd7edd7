"""Rows a1-a3 on the GPU: batched intersect_with / subtracted_by /
intersect_and_count / trim_v_list_lower_bound against the oracle's
restatement of src/vertex_set.cc and include/utilities.h, bit-exact, on
random and adversarial lists (empty, disjoint, identical, 1-vs-100K skew,
lists longer than the LDS hash, ids near 2^32-1)."""

import random

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


def _lists():
    rng = random.Random(7)
    top = 2**32 - 2
    L = [[], [5], list(range(0, 64)), list(range(1, 65, 2)), list(range(0, 200000, 2)),
         [top - 3, top - 1, top], [top - 1], list(range(100000)), list(range(3000, 9000)),
         list(range(0, 9000, 3))]
    for _ in range(300):
        hi = rng.choice([100, 5000, 2**32 - 2])
        L.append(sorted(rng.sample(range(hi), min(hi, rng.choice([0, 1, 3, 30, 200, 1500, 5000])))))
    return L


def test_set_primitives_match_oracle(kzdev):
    kz = kzdev
    lists = _lists()
    rng = random.Random(3)
    pairs = [(i, j) for i in range(10) for j in range(10)] + \
            [(rng.randrange(len(lists)), rng.randrange(len(lists))) for _ in range(3000)]
    inter = kz.set_op_batch("intersect", lists, pairs)
    sub = kz.set_op_batch("subtract", lists, pairs)
    cnt = kz.set_op_batch("intersect_count", lists, pairs)
    for (a, b), x, y, c in zip(pairs, inter, sub, cnt):
        A, B = lists[a], lists[b]
        assert x.tolist() == oracle.intersect(A, B).tolist()
        assert y.tolist() == oracle.subtract(A, B).tolist()
        assert int(c) == oracle.intersect_count(A, B)
    bounds = [rng.choice([0, 1, 5, 63, 64, 4999, 2**32 - 2, 2**32 - 1]) for _ in pairs]
    tr = kz.set_op_batch("trim_lower_bound", lists, pairs, bounds=bounds)
    for (a, _), bd, t in zip(pairs, bounds, tr):
        assert int(t) == oracle.trim_lower_bound(lists[a], bd)


def test_set_primitives_reject_small_output(kzdev):
    import ctypes as C

    from openkhuzdulgraph_amd import _ffi

    vals = np.array([1, 2, 3, 2, 3], np.uint32)
    off = np.array([0, 3, 5], np.uint64)
    a = np.array([0], np.uint32)
    b = np.array([1], np.uint32)
    oo = np.array([0, 1], np.uint64)  # needs 2
    out = np.zeros(2, np.uint32)
    ln = np.zeros(1, np.uint32)
    P32, P64 = C.POINTER(C.c_uint32), C.POINTER(C.c_uint64)
    rc = _ffi.lib().kz_set_op_batch(_ffi.KZ_SETOP_INTERSECT, vals.ctypes.data_as(P32), off.ctypes.data_as(P64), 2,
                                    a.ctypes.data_as(P32), b.ctypes.data_as(P32), None, 1,
                                    out.ctypes.data_as(P32), oo.ctypes.data_as(P64), ln.ctypes.data_as(P32))
    assert rc == _ffi.KZ_ERR_INVALID
