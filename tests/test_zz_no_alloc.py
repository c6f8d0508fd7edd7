"""kz_set_op_batch_noalloc (no device allocation; staging in a file-scope
device array) against kz_set_op_batch, through the C ABI: the same results,
bit for bit, for every operation on two calls in a row, the second with a
smaller batch (the array keeps the first call's contents); a batch beyond
the array's size is refused before any device work."""

import random

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _batch(rng, nlists, npairs, maxlen):
    lists = []
    for _ in range(nlists):
        hi = rng.choice([64, 5000, 2**32 - 2])
        n = min(hi, rng.choice([0, 1, 3, 40, 300, maxlen]))
        lists.append(sorted(rng.sample(range(hi), n)))
    pairs = [(rng.randrange(nlists), rng.randrange(nlists)) for _ in range(npairs)]
    bounds = [rng.choice([0, 7, 63, 4999, 2**32 - 2]) for _ in range(npairs)]
    return lists, pairs, bounds


def _same(a, b):
    if isinstance(a, list):
        return len(a) == len(b) and all(x.tolist() == y.tolist() for x, y in zip(a, b))
    return a.tolist() == b.tolist()


def test_noalloc_matches_the_allocating_entry_point(kzdev):
    kz = kzdev
    rng = random.Random(11)
    big = _batch(rng, 400, 6000, 3000)
    small = _batch(rng, 30, 200, 500)
    for name, (lists, pairs, bounds) in (("large", big), ("smaller", small)):
        for op in ("intersect", "subtract", "intersect_count", "trim_lower_bound"):
            print(f"[case] no-alloc {name} op={op}", flush=True)
            bd = bounds if op == "trim_lower_bound" else None
            want = kz.set_op_batch(op, lists, pairs, bounds=bd)
            got = kz.set_op_batch(op, lists, pairs, bounds=bd, noalloc=True)
            assert _same(got, want), (name, op)


def test_noalloc_refuses_a_batch_beyond_its_array(kzdev):
    kz = kzdev
    from openkhuzdulgraph_amd import _ffi
    n = _ffi.KZ_SETOP_NOALLOC_MAX_BYTES // 4 + 1
    lists = [np.arange(n, dtype=np.uint32)]
    with pytest.raises(kz.KzError) as ei:
        kz.set_op_batch("intersect_count", lists, [(0, 0)], noalloc=True)
    assert ei.value.code == _ffi.KZ_ERR_INVALID
    # the allocating entry point takes the same batch
    assert int(kz.set_op_batch("intersect_count", lists, [(0, 0)])[0]) == n
