// Serving: a worker thread runs queries on resident graphs (kz_server_*).
//
// No reference counterpart: the reference runs one app (one pattern, one
// graph) per MPI job, DistributedApplication::run
// (src/distributed_application.cc:259-356).  Here many queries share one
// device:
//   * admission: queries are taken in arrival order; a query needs a
//     prepared execution (a clique job's task lists and slabs, or the census
//     copy); the bytes of all prepared executions the server keeps stay
//     within the budget, idle ones are released least recently used first;
//   * continuous batching: an admitted clique query is launched on a free
//     stream as soon as its execution is ready (up to max_inflight at once),
//     so new queries start while earlier ones still run; completion is
//     polled with events and results land in pinned host memory;
//   * a census query runs on the worker thread itself (its passes are
//     ordered on the library stream) while launched clique queries proceed
//     on their streams.

#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "kz_device.h"
#include "kz_kernels.h"

namespace kz {
namespace {

using Clock = std::chrono::steady_clock;

double ms_between(Clock::time_point a, Clock::time_point b) {
    return std::chrono::duration<double, std::milli>(b - a).count();
}

struct SQuery {
    uint64_t ticket = 0;
    kz_graph g = nullptr;
    int kind = 0, k = 0;
    Clock::time_point t_submit, t_start;
    kz_query_result res{};
    bool done = false;
};

// A prepared execution whose device memory the server accounts for.
struct Prepared {
    kz_graph g = nullptr;
    int kind = 0, k = 0;
    uint64_t bytes = 0;
    uint64_t last_use = 0;
    int busy = 0;  // launched queries using it
};

}  // namespace
}  // namespace kz

struct kz_server_s {
    kz_server_cfg cfg{};
    int device = 0;
    std::thread worker;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<kz::SQuery*> pending;
    std::unordered_map<uint64_t, kz::SQuery*> live;  // submitted, not yet retired
    uint64_t next_ticket = 1;
    bool stopping = false;
    // worker-only state
    std::vector<hipStream_t> streams;
    std::vector<hipEvent_t> events;
    std::vector<kz::SQuery*> slot_query;
    std::vector<kz::Prepared*> slot_prep;
    unsigned long long* pinned = nullptr;  // one result per stream slot
    std::vector<kz::Prepared*> prepared;
    uint64_t in_use = 0, tick = 0;
};

namespace kz {
namespace {

uint64_t budget_of(const kz_server_s* s) { return s->cfg.hbm_budget ? s->cfg.hbm_budget : ~uint64_t(0); }

Prepared* find_prepared(kz_server_s* s, kz_graph g, int kind, int k) {
    for (Prepared* p : s->prepared)
        if (p->g == g && p->kind == kind && p->k == k) return p;
    return nullptr;
}

// Kind of the Prepared entry that accounts for the forward copy of a
// symmetric graph, which every clique execution on that graph runs on.
constexpr int kForwardCopy = -1;

bool clique_exec_on(const DeviceGraph& d) {
    for (int k = 3; k <= 5; ++k)
        if (d.clique_exec[k]) return true;
    return false;
}

void release_prepared(kz_server_s* s, Prepared* p) {
    DeviceGraph& d = p->g->g;
    if (p->kind == KZ_QUERY_CLIQUE) d.clique_exec[p->k].reset();
    else if (p->kind == kForwardCopy) {
        // still in use by a clique execution (one being prepared): freed
        // with it
        if (!clique_exec_on(d)) release_forward(d);
    } else {
        d.census.reset();
    }
    s->in_use -= p->bytes;
    for (size_t i = 0; i < s->prepared.size(); ++i)
        if (s->prepared[i] == p) {
            s->prepared.erase(s->prepared.begin() + long(i));
            break;
        }
    delete p;
}

// Device bytes the degree orientation of a symmetric graph takes while it
// is built: ranks and sort keys, the forward lists (half the entries) and
// the keys of the bucketed sort (16 bytes per forward entry at most).
uint64_t forward_estimate(const DeviceGraph& d) {
    return 8ull * (d.n + 1) + 28ull * d.n + col_alloc_bytes(d.nnz) + 8ull * d.nnz + (1u << 20);
}

// Device bytes a not yet prepared execution will take (upper estimates).
uint64_t estimate_bytes(const DeviceGraph& d, int kind, int k) {
    if (kind == KZ_QUERY_CLIQUE) {
        // task lists and cursors; for k >= 4 the hub roots' bitmap rows in
        // HBM (d * ceil(d / 64) words each, out-degrees above the default
        // 1024 bound); a symmetric graph's forward copy is accounted apart
        const std::vector<uint32_t>& fd = d.oriented ? d.h_deg : d.h_fdeg;
        uint64_t hubs = 0;
        if (k >= 4)
            for (uint32_t x : fd)
                if (x > 1024) hubs += uint64_t(x) * ((x + 63) / 64) * 8;
        return 8ull * d.n + (1u << 20) + hubs;
    }
    // relabelled copy, forward part, edge counters, 4-cycle slabs (at most
    // 2 workgroups per CU) and the 4-clique job
    return 8ull * (d.n + 1) * 3 + col_alloc_bytes(d.nnz) * 3 + 16ull * d.n * uint64_t(rt().num_cus) +
           8ull * d.n + (1u << 20);
}

// Device memory kept free next to what the server prepares.
constexpr uint64_t kFreeMargin = uint64_t(1) << 30;

uint64_t free_device_bytes() {
    size_t f = 0, t = 0;
    if (hipMemGetInfo(&f, &t) != hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    return f;
}

// An idle execution can be released; a forward copy once no clique
// execution on its graph is left.
bool releasable(const Prepared* p) {
    if (p->busy) return false;
    return p->kind != kForwardCopy || !clique_exec_on(p->g->g);
}

// Frees idle prepared executions, least recently used first, until `need`
// more bytes fit in the budget and in the device's free memory.  Returns
// whether they fit.
bool make_room(kz_server_s* s, uint64_t need) {
    const uint64_t B = budget_of(s);
    while (s->in_use + need > B || (need && free_device_bytes() < need + kFreeMargin)) {
        Prepared* victim = nullptr;
        for (Prepared* p : s->prepared)
            if (releasable(p) && (!victim || p->last_use < victim->last_use)) victim = p;
        if (!victim) return false;
        release_prepared(s, victim);
    }
    return true;
}

void finish(kz_server_s* s, SQuery* q, int status) {
    q->res.status = status;
    const Clock::time_point t = Clock::now();
    q->res.total_ms = ms_between(q->t_submit, t);
    if (q->t_start == Clock::time_point{}) q->t_start = t;
    q->res.queue_ms = ms_between(q->t_submit, q->t_start);
    q->res.run_ms = ms_between(q->t_start, t);
    std::lock_guard<std::mutex> lk(s->mu);
    q->done = true;
    s->cv_done.notify_all();
}

// Releases every idle prepared execution except `keep`; returns whether any
// was released.
bool release_idle(kz_server_s* s, Prepared* keep) {
    bool any = false;
    for (bool again = true; again;) {
        again = false;
        for (Prepared* p : s->prepared)
            if (p != keep && releasable(p)) {
                release_prepared(s, p);
                any = again = true;
                break;
            }
    }
    return any;
}

bool any_busy(const kz_server_s* s) {
    for (const Prepared* p : s->prepared)
        if (p->busy) return true;
    return false;
}

// Accounts a fresh preparation of `bytes`; trims idle ones if the estimate
// was low.  Returns false (and releases it) when it still does not fit.
bool account(kz_server_s* s, Prepared* p, uint64_t bytes) {
    p->bytes = bytes;
    s->in_use += bytes;
    p->busy++;  // keep it while trimming the others
    const bool fits = make_room(s, 0);
    p->busy--;
    if (!fits) release_prepared(s, p);
    return fits;
}

// Accounts the forward copy of a symmetric graph (built by graph_job) once
// per graph.  Returns false when it does not fit the budget; the entry is
// then dropped.
bool adopt_forward(kz_server_s* s, kz_graph g) {
    DeviceGraph& d = g->g;
    if (d.oriented || !d.fwd_ready || !d.fwd_bytes) return true;
    if (Prepared* f = find_prepared(s, g, kForwardCopy, 0)) {
        f->last_use = ++s->tick;
        return true;
    }
    Prepared* f = new Prepared{g, kForwardCopy, 0, 0, ++s->tick, 0};
    s->prepared.push_back(f);
    return account(s, f, d.fwd_bytes);
}

// Drops a clique execution that could not be accounted, and the graph's
// forward copy when nothing uses or accounts for it any more.
void drop_clique(kz_server_s* s, Prepared* p) {
    const kz_graph g = p->g;
    release_prepared(s, p);
    DeviceGraph& d = g->g;
    if (d.fwd_ready && !clique_exec_on(d) && !find_prepared(s, g, kForwardCopy, 0)) release_forward(d);
}

// Prepares (if needed) and launches or runs one query.  Returns false when
// it must wait for memory held by running queries.
bool start_query(kz_server_s* s, SQuery* q, int slot) {
    DeviceGraph& d = q->g->g;
    Prepared* p = find_prepared(s, q->g, q->kind, q->k);
    if (!p) {
        // an execution prepared before (by a direct call or an earlier
        // server) is adopted with its real size
        const bool cached = q->kind == KZ_QUERY_CLIQUE ? bool(d.clique_exec[q->k]) : bool(d.census);
        if (cached) {
            p = new Prepared{q->g, q->kind, q->k, 0, 0, 0};
            s->prepared.push_back(p);
            const uint64_t bytes =
                q->kind == KZ_QUERY_CLIQUE ? static_cast<CliqueJob*>(d.clique_exec[q->k].get())->bytes : census_bytes(d);
            p->busy++;
            const bool fwd_fits = q->kind != KZ_QUERY_CLIQUE || adopt_forward(s, q->g);
            p->busy--;
            if (!fwd_fits) drop_clique(s, p);
            if (!fwd_fits || !account(s, p, bytes)) {
                if (!any_busy(s)) finish(s, q, KZ_ERR_BUDGET);
                return !any_busy(s);
            }
        }
    }
    if (!p && q->kind == KZ_QUERY_CLIQUE && !d.oriented && !d.fwd_ready) {
        // a symmetric graph's first clique query: its degree orientation
        // first (accounted with the graph), then the job, sized from the
        // exact forward degrees
        const uint64_t est = forward_estimate(d);
        if (est > budget_of(s)) {
            finish(s, q, KZ_ERR_BUDGET);
            return true;
        }
        if (!make_room(s, est)) {
            if (!any_busy(s)) finish(s, q, KZ_ERR_BUDGET);
            return !any_busy(s);
        }
        int e = ensure_forward(d);
        if ((e == KZ_ERR_DEVICE || e == KZ_ERR_NOMEM) && release_idle(s, nullptr)) e = ensure_forward(d);
        if (e) {
            finish(s, q, e);
            return true;
        }
        if (!adopt_forward(s, q->g)) {
            release_forward(d);
            if (!any_busy(s)) finish(s, q, KZ_ERR_BUDGET);
            return !any_busy(s);
        }
    }
    if (!p) {
        const uint64_t est = estimate_bytes(d, q->kind, q->k);
        Prepared* fwd = find_prepared(s, q->g, kForwardCopy, 0);
        if (fwd) fwd->busy++;  // the job needs it: not a victim for its room
        const bool room = est <= budget_of(s) && make_room(s, est);
        if (fwd) fwd->busy--;
        if (est > budget_of(s)) {
            finish(s, q, KZ_ERR_BUDGET);
            return true;
        }
        if (!room) {
            // nothing running would free more: it cannot fit
            if (!any_busy(s)) finish(s, q, KZ_ERR_BUDGET);
            return !any_busy(s);
        }
        p = new Prepared{q->g, q->kind, q->k, 0, 0, 0};
        s->prepared.push_back(p);
        if (q->kind == KZ_QUERY_CLIQUE) {
            CliqueJob* job = nullptr;
            int e = graph_job(d, q->k, &job);
            // out of device memory despite the estimate: release every idle
            // preparation and try once more
            if ((e == KZ_ERR_DEVICE || e == KZ_ERR_NOMEM) && release_idle(s, p)) e = graph_job(d, q->k, &job);
            if (e) {
                release_prepared(s, p);
                finish(s, q, e);
                return true;
            }
            // the forward copy the job runs on belongs to the graph and is
            // accounted once for all its clique executions
            p->busy++;
            const bool fwd_fits = adopt_forward(s, q->g);
            p->busy--;
            if (!fwd_fits) drop_clique(s, p);
            if (!fwd_fits || !account(s, p, job->bytes)) {
                if (!any_busy(s)) finish(s, q, KZ_ERR_BUDGET);
                return !any_busy(s);
            }
        }
    }
    p->last_use = ++s->tick;
    q->t_start = Clock::now();
    if (q->kind == KZ_QUERY_CENSUS) {
        // runs here, ordered on the library stream
        kz_run_stats st;
        const int e = motif_census(d, q->res.counts, &st);
        if (!e && p->bytes == 0 && !account(s, p, census_bytes(d))) {
            // it ran, but its state does not fit next to the busy ones: drop it
        }
        finish(s, q, e);
        return true;
    }
    CliqueJob* job = nullptr;
    if (int e = graph_job(d, q->k, &job)) {
        finish(s, q, e);
        return true;
    }
    p->busy++;
    s->slot_query[slot] = q;
    s->slot_prep[slot] = p;
    if (clique_job_launch(*job, s->streams[slot], s->pinned + slot) != KZ_OK ||
        hipEventRecord(s->events[slot], s->streams[slot]) != hipSuccess) {
        (void)hipGetLastError();
        p->busy--;
        s->slot_query[slot] = nullptr;
        s->slot_prep[slot] = nullptr;
        finish(s, q, KZ_ERR_DEVICE);
    }
    return true;
}

}  // namespace
}  // namespace kz

using namespace kz;

static void worker_main(kz_server_s* s) {
    (void)hipSetDevice(s->device);
    const int nslots = int(s->streams.size());
    while (true) {
        // completions
        int busy_slots = 0;
        for (int i = 0; i < nslots; ++i) {
            SQuery* q = s->slot_query[i];
            if (!q) continue;
            const hipError_t e = hipEventQuery(s->events[i]);
            if (e == hipErrorNotReady) {
                ++busy_slots;
                continue;
            }
            if (e == hipSuccess) q->res.counts[0] = s->pinned[i];
            else (void)hipGetLastError();
            s->slot_prep[i]->busy--;
            s->slot_query[i] = nullptr;
            s->slot_prep[i] = nullptr;
            finish(s, q, e == hipSuccess ? KZ_OK : KZ_ERR_DEVICE);
        }
        // admission and launch, in arrival order; a query whose execution is
        // in use by a running query waits for it, later ones may go first
        std::vector<SQuery*> batch;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            if (s->pending.empty() && busy_slots == 0) {
                if (s->stopping) break;
                s->cv_work.wait(lk, [&] { return s->stopping || !s->pending.empty(); });
                continue;
            }
            batch.assign(s->pending.begin(), s->pending.end());
        }
        std::vector<SQuery*> started;
        for (SQuery* q : batch) {
            Prepared* p = find_prepared(s, q->g, q->kind, q->k);
            if (p && p->busy) continue;
            int slot = -1;
            if (q->kind == KZ_QUERY_CLIQUE) {
                for (int i = 0; i < nslots; ++i)
                    if (!s->slot_query[i]) {
                        slot = i;
                        break;
                    }
                if (slot < 0) break;
            }
            // the memory check must not overtake an earlier waiting query
            if (!start_query(s, q, slot)) break;
            started.push_back(q);
            ++busy_slots;
        }
        if (!started.empty()) {
            std::lock_guard<std::mutex> lk(s->mu);
            for (SQuery* q : started)
                for (auto it = s->pending.begin(); it != s->pending.end(); ++it)
                    if (*it == q) {
                        s->pending.erase(it);
                        break;
                    }
        } else if (busy_slots) {
            std::unique_lock<std::mutex> lk(s->mu);
            s->cv_work.wait_for(lk, std::chrono::microseconds(50));
        }
    }
}

extern "C" {

int kz_server_start(const kz_server_cfg* cfg, kz_server* out) {
    if (!cfg || !out) KZ_FAIL(KZ_ERR_INVALID, "null argument");
    if (cfg->max_inflight < 1 || cfg->max_inflight > 16) KZ_FAIL(KZ_ERR_INVALID, "max_inflight must be 1..16");
    if (int e = ensure_ready()) return e;
    kz_server_s* s = new (std::nothrow) kz_server_s();
    if (!s) KZ_FAIL(KZ_ERR_NOMEM, "server");
    s->cfg = *cfg;
    s->device = rt().device;
    const int n = int(cfg->max_inflight);
    s->streams.assign(n, nullptr);
    s->events.assign(n, nullptr);
    s->slot_query.assign(n, nullptr);
    s->slot_prep.assign(n, nullptr);
    auto fail = [&]() {
        for (auto st : s->streams)
            if (st) (void)hipStreamDestroy(st);
        for (auto ev : s->events)
            if (ev) (void)hipEventDestroy(ev);
        if (s->pinned) (void)hipHostFree(s->pinned);
        delete s;
        (void)hipGetLastError();
        KZ_FAIL(KZ_ERR_DEVICE, "server streams, events or pinned memory");
    };
    for (int i = 0; i < n; ++i)
        if (hipStreamCreateWithFlags(&s->streams[i], hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&s->events[i], hipEventDisableTiming) != hipSuccess)
            return fail();
    if (hipHostMalloc(reinterpret_cast<void**>(&s->pinned), sizeof(unsigned long long) * n, hipHostMallocDefault) !=
        hipSuccess)
        return fail();
    s->worker = std::thread(worker_main, s);
    *out = s;
    return KZ_OK;
}

int kz_server_submit(kz_server s, kz_graph g, int kind, int k, uint64_t* ticket) {
    if (!s || !g || !ticket) KZ_FAIL(KZ_ERR_INVALID, "null argument");
    if (kind == KZ_QUERY_CLIQUE) {
        if (k < 3 || k > 5) KZ_FAIL(KZ_ERR_UNSUPPORTED, "clique size %d outside 3..5", k);
    } else if (kind == KZ_QUERY_CENSUS) {
        if (g->g.oriented) KZ_FAIL(KZ_ERR_INVALID, "the census needs a symmetric graph");
        k = 0;
    } else {
        KZ_FAIL(KZ_ERR_INVALID, "unknown query kind %d", kind);
    }
    SQuery* q = new (std::nothrow) SQuery();
    if (!q) KZ_FAIL(KZ_ERR_NOMEM, "query");
    q->g = g;
    q->kind = kind;
    q->k = k;
    q->t_submit = Clock::now();
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (s->stopping) {
            delete q;
            KZ_FAIL(KZ_ERR_STATE, "server is stopping");
        }
        q->ticket = s->next_ticket++;
        s->live[q->ticket] = q;
        s->pending.push_back(q);
    }
    s->cv_work.notify_one();
    *ticket = q->ticket;
    return KZ_OK;
}

int kz_server_poll(kz_server s, uint64_t ticket, int wait, kz_query_result* out) {
    if (!s || !out) KZ_FAIL(KZ_ERR_INVALID, "null argument");
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->live.find(ticket);
    if (it == s->live.end()) KZ_FAIL(KZ_ERR_INVALID, "unknown or retired ticket %llu", (unsigned long long)ticket);
    SQuery* q = it->second;
    if (!q->done) {
        if (!wait) return KZ_PENDING;
        s->cv_done.wait(lk, [&] { return q->done; });
    }
    *out = q->res;
    s->live.erase(ticket);
    delete q;
    return KZ_OK;
}

int kz_server_stop(kz_server s) {
    if (!s) KZ_FAIL(KZ_ERR_INVALID, "null argument");
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stopping = true;
    }
    s->cv_work.notify_all();
    if (s->worker.joinable()) s->worker.join();
    for (auto st : s->streams) (void)hipStreamDestroy(st);
    for (auto ev : s->events) (void)hipEventDestroy(ev);
    (void)hipHostFree(s->pinned);
    // the executions prepared under this server's budget are released
    while (!s->prepared.empty()) release_prepared(s, s->prepared.back());
    for (auto& kv : s->live) delete kv.second;
    delete s;
    return KZ_OK;
}

}  // extern "C"
