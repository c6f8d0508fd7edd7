"""GPU tests of the query server (kz_server_*): every query in a mixed,
concurrently running stream returns the same exact count as the oracle
(alone-vs-mix parity), under no budget, under a budget that forces prepared
executions to be evicted and re-prepared, and with a budget too small for
any query (KZ_ERR_BUDGET)."""

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

APP = {3: "kautomine_triangle", 4: "kautomine_four_clique", 5: "kautomine_five_clique"}


def _pool(kz):
    sym = {
        "chunglu": kz.gen_chung_lu(6000, 30000, 2.3, 400, 3),
        "rmat11": kz.gen_rmat(11, 12, seed=9),
    }
    ori = {
        "rmat12_o": kz.gen_rmat(12, 16, seed=4).oriented(),
        "er_o": kz.gen_erdos_renyi(3000, 45000, 6).oriented(),
        "chunglu_o": sym["chunglu"].oriented(),
    }
    return sym, ori


def _expected(sym, ori):
    want = {}
    for name, g in ori.items():
        for k in (3, 4, 5):
            want[(name, "clique", k)] = [oracle.run_app(APP[k], g.rowptr, g.col)[1][0]]
    for name, g in sym.items():
        want[(name, "census", 0)] = oracle.motif_census(g.rowptr, g.col)
        want[(name, "clique", 4)] = [oracle.run_app(APP[4], g.rowptr, g.col)[1][0]]
    return want


def _run_stream(kz, server, dgs, queries):
    tickets = [(key, server.submit(dgs[key[0]], key[1], key[2] or 4)) for key in queries]
    out = []
    for key, t in tickets:
        res = server.poll(t, wait=True)
        out.append((key, res))
    return out


def test_server_mixed_stream_exact(kzdev):
    kz = kzdev
    sym, ori = _pool(kz)
    want = _expected(sym, ori)
    dgs = {name: g.to_device() for name, g in {**sym, **ori}.items()}
    keys = sorted(want)
    rng = np.random.default_rng(5)
    queries = [keys[i] for i in rng.integers(0, len(keys), 60)]
    print(f"[case] server mixed stream, {len(queries)} queries over {len(keys)} kinds", flush=True)
    with kz.Server(hbm_budget=0, max_inflight=4) as srv:
        for key, res in _run_stream(kz, srv, dgs, queries):
            assert res["status"] == 0, (key, res)
            n = len(want[key])
            assert res["counts"][:n] == want[key], (key, res["counts"], want[key])
            assert res["total_ms"] >= res["run_ms"] >= 0
        # non-blocking poll eventually returns the result
        t = srv.submit(dgs["rmat12_o"], "clique", 4)
        res = None
        while res is None:
            res = srv.poll(t)
        assert res["counts"][0] == want[("rmat12_o", "clique", 4)][0]
        with pytest.raises(kz.KzError):
            srv.poll(t)  # retired


def test_server_budget_eviction_and_rejection(kzdev):
    kz = kzdev
    sym, ori = _pool(kz)
    want = _expected(sym, ori)
    dgs = {name: g.to_device() for name, g in {**sym, **ori}.items()}
    keys = [k for k in sorted(want) if k[1] == "clique"]
    # room for roughly one prepared clique execution at a time: every switch
    # of (graph, k) evicts the previous one, counts must stay exact
    budget = 8 * 6000 + (1 << 20) + 8 * 6001 + 4 * (2 * 30000 + 8)
    print(f"[case] server eviction, budget {budget}", flush=True)
    with kz.Server(hbm_budget=budget * 3, max_inflight=2) as srv:
        for key, res in _run_stream(kz, srv, dgs, keys * 2):
            assert res["status"] == 0, (key, res)
            assert res["counts"][0] == want[key][0], key
    print("[case] server rejection", flush=True)
    with kz.Server(hbm_budget=4096, max_inflight=2) as srv:
        for key, res in _run_stream(kz, srv, dgs, keys[:3] + [("chunglu", "census", 0)]):
            assert res["status"] == kz._ffi.KZ_ERR_BUDGET, (key, res)


def test_queries_are_admitted_during_a_census_and_a_preparation(kzdev):
    """A census runs on a helper thread, and a clique job is prepared on
    another: a small clique query submitted behind either is admitted,
    launched and finished while they are still running, and every result
    stays exact."""
    kz = kzdev
    big_sym = kz.gen_chung_lu(1500000, 12000000, 2.3, 8000, 7)
    big_ori = kz.gen_rmat(21, 16, seed=3).oriented()
    small = kz.gen_rmat(12, 16, seed=4).oriented()
    want_small = oracle.run_app(APP[4], small.rowptr, small.col)[1][0]
    want_census = oracle.motif_census(big_sym.rowptr, big_sym.col)
    want_big = oracle.run_app(APP[3], big_ori.rowptr, big_ori.col)[1][0]
    d_sym, d_ori, d_small = big_sym.to_device(), big_ori.to_device(), small.to_device()
    with kz.Server(hbm_budget=0, max_inflight=4) as srv:
        # the small query's execution is prepared before the timed part
        assert srv.poll(srv.submit(d_small, "clique", 4), wait=True)["counts"][0] == want_small
        print("[case] clique query behind a census", flush=True)
        tc = srv.submit(d_sym, "census", 0)
        tq = srv.submit(d_small, "clique", 4)
        rq = srv.poll(tq, wait=True)
        census_running = srv.poll(tc) is None
        rc = srv.poll(tc, wait=True)
        assert rq["status"] == 0 and rq["counts"][0] == want_small
        assert rc["status"] == 0 and rc["counts"] == want_census
        assert census_running, (rq, rc)
        print("[case] clique query behind a preparation", flush=True)
        tb = srv.submit(d_ori, "clique", 3)
        tq = srv.submit(d_small, "clique", 4)
        rq = srv.poll(tq, wait=True)
        big_running = srv.poll(tb) is None
        rb = srv.poll(tb, wait=True)
        assert rq["status"] == 0 and rq["counts"][0] == want_small
        assert rb["status"] == 0 and rb["counts"][0] == want_big
        assert big_running, (rq, rb)
