// Serving: queries on resident graphs (kz_server_*).
//
// No reference counterpart: the reference runs one app (one pattern, one
// graph) per MPI job, DistributedApplication::run
// (src/distributed_application.cc:259-356).  Here many queries share one
// device.  A worker thread owns the admission state; two helper threads do
// the slow host-side work so that it never holds up admission:
//   * admission: queries are taken in arrival order; a query needs a
//     prepared execution (a clique job's task lists and slabs, a symmetric
//     graph's forward copy, or the census copy); the bytes of every
//     prepared execution the server keeps (and the estimates of those being
//     prepared) stay within the budget, idle ones are released cheapest to
//     rebuild first (GreedyDual over their measured preparation times, so a
//     large graph's costly job outlives many small ones); a query that must
//     wait for memory held by running queries holds back the ones behind it;
//   * preparation (helper thread 1): forward copies and clique jobs are built
//     while other queries keep being admitted and launched; a query whose
//     execution is being prepared is passed by later ones;
//   * continuous batching: an admitted clique query is launched on a free
//     stream as soon as its execution is ready (up to max_inflight at once);
//     completions are polled with events, results land in pinned memory;
//   * census queries (helper thread 2) run beside the clique queries.

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
    std::string err;  // kz_last_error() of the call that failed the query
    bool retried = false;  // a census that failed with a device error runs once more
    bool done = false;
};

// Kind of the Prepared entry that accounts for the forward copy of a
// symmetric graph, which every clique execution on that graph runs on.
constexpr int kForwardCopy = -1;

// A prepared execution whose device memory the server accounts for.
struct Prepared {
    kz_graph g = nullptr;
    int kind = 0, k = 0;
    uint64_t bytes = 0;  // accounted (an estimate while preparing)
    double cost_ms = 0;  // measured preparation time
    double H = 0;        // GreedyDual priority: released lowest first
    int busy = 0;        // launched queries or helper work using it
    bool preparing = false;
    bool retried = false;
};

// Work for a helper thread.
enum { kJobForward = 1, kJobPrepare = 2, kJobCensus = 3 };
struct AuxJob {
    int type = 0;
    SQuery* q = nullptr;
    Prepared* p = nullptr;
    int rc = 0;
    double ms = 0;
    std::string err;  // the helper thread's kz_last_error() when rc != 0
};

struct AuxWorker {
    std::thread th;
    std::mutex mu;
    std::condition_variable cv;
    std::deque<AuxJob> todo;
    std::vector<AuxJob> done;
    bool stop = false;
    int pending = 0;  // submitted, not yet collected (worker thread only)
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
    bool aux_event = false;  // a helper finished a job (guarded by mu)
    // worker-only state
    std::vector<hipStream_t> streams;
    std::vector<hipEvent_t> events;
    std::vector<kz::SQuery*> slot_query;
    std::vector<kz::Prepared*> slot_prep;
    unsigned long long* pinned = nullptr;  // one result per stream slot
    std::vector<kz::Prepared*> prepared;
    uint64_t in_use = 0;
    double L = 0;  // GreedyDual inflation
    kz::AuxWorker prep, census;
};

namespace kz {
namespace {

uint64_t budget_of(const kz_server_s* s) { return s->cfg.hbm_budget ? s->cfg.hbm_budget : ~uint64_t(0); }

Prepared* find_prepared(kz_server_s* s, kz_graph g, int kind, int k) {
    for (Prepared* p : s->prepared)
        if (p->g == g && p->kind == kind && p->k == k) return p;
    return nullptr;
}

bool clique_exec_on(const DeviceGraph& d) {
    for (int k = 3; k <= 5; ++k)
        if (d.clique_exec[k]) return true;
    return false;
}

// A clique execution of graph g is kept or being prepared.  Checked before
// the graph's own slots are read: a helper may be writing one of them.
bool clique_prepared_on(kz_server_s* s, kz_graph g) {
    for (Prepared* p : s->prepared)
        if (p->g == g && p->kind == KZ_QUERY_CLIQUE) return true;
    return false;
}

void touch(kz_server_s* s, Prepared* p) { p->H = s->L + p->cost_ms + 1e-3; }

void release_prepared(kz_server_s* s, Prepared* p) {
    DeviceGraph& d = p->g->g;
    if (p->kind == KZ_QUERY_CLIQUE) {
        d.clique_exec[p->k].reset();
    } else if (p->kind == kForwardCopy) {
        // still used by a clique execution kept outside the server: freed
        // with the graph
        if (!clique_prepared_on(s, p->g) && !clique_exec_on(d)) release_forward(d);
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

// An idle execution can be released; a forward copy once no clique
// execution on its graph is kept or being prepared.
bool releasable(kz_server_s* s, const Prepared* p) {
    if (p->busy || p->preparing) return false;
    return p->kind != kForwardCopy || (!clique_prepared_on(s, p->g) && !clique_exec_on(p->g->g));
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

// Frees idle prepared executions, lowest GreedyDual priority first, until
// `need` more bytes fit in the budget and in the device's free memory.
// Returns whether they fit.
bool make_room(kz_server_s* s, uint64_t need) {
    const uint64_t B = budget_of(s);
    while (s->in_use + need > B || (need && free_device_bytes() < need + kFreeMargin)) {
        Prepared* victim = nullptr;
        for (Prepared* p : s->prepared)
            if (releasable(s, p) && (!victim || p->H < victim->H)) victim = p;
        if (!victim) return false;
        s->L = victim->H;
        release_prepared(s, victim);
    }
    return true;
}

// Releases every idle prepared execution except `keep`; returns whether any
// was released.
bool release_idle(kz_server_s* s, Prepared* keep) {
    bool any = false;
    for (bool again = true; again;) {
        again = false;
        for (Prepared* p : s->prepared)
            if (p != keep && releasable(s, p)) {
                release_prepared(s, p);
                any = again = true;
                break;
            }
    }
    return any;
}

bool any_busy(const kz_server_s* s) {
    for (const Prepared* p : s->prepared)
        if (p->busy || p->preparing) return true;
    return false;
}

// err: the message of a failure on another thread (a helper's), else the
// calling thread's own kz_last_error() is kept.
void finish(kz_server_s* s, SQuery* q, int status, const std::string& err = std::string()) {
    q->res.status = status;
    if (status) q->err = err.empty() ? std::string(get_error()) : err;
    const Clock::time_point t = Clock::now();
    q->res.total_ms = ms_between(q->t_submit, t);
    if (q->t_start == Clock::time_point{}) q->t_start = t;
    q->res.queue_ms = ms_between(q->t_submit, q->t_start);
    q->res.run_ms = ms_between(q->t_start, t);
    std::lock_guard<std::mutex> lk(s->mu);
    q->done = true;
    s->cv_done.notify_all();
}

// ------------------------------------------------------- helper threads --

void aux_run(AuxJob& j) {
    const Clock::time_point t0 = Clock::now();
    DeviceGraph& d = j.p->g->g;
    if (j.type == kJobForward) {
        j.rc = ensure_forward(d);
    } else if (j.type == kJobPrepare) {
        CliqueJob* job = nullptr;
        j.rc = graph_job(d, j.p->k, &job);
    } else {
        kz_run_stats st;
        j.q->t_start = t0;
        j.rc = motif_census(d, j.q->res.counts, &st);
    }
    if (j.rc) j.err = get_error();
    j.ms = ms_between(t0, Clock::now());
}

void aux_loop(kz_server_s* s, AuxWorker* w) {
    (void)hipSetDevice(s->device);
    while (true) {
        AuxJob j;
        {
            std::unique_lock<std::mutex> lk(w->mu);
            w->cv.wait(lk, [&] { return w->stop || !w->todo.empty(); });
            if (w->todo.empty()) return;
            j = w->todo.front();
            w->todo.pop_front();
        }
        aux_run(j);
        {
            std::lock_guard<std::mutex> lk(w->mu);
            w->done.push_back(j);
        }
        {
            std::lock_guard<std::mutex> lk(s->mu);
            s->aux_event = true;
        }
        s->cv_work.notify_one();
    }
}

void aux_submit(AuxWorker& w, const AuxJob& j) {
    {
        std::lock_guard<std::mutex> lk(w.mu);
        w.todo.push_back(j);
    }
    w.pending++;
    w.cv.notify_one();
}

std::vector<AuxJob> aux_collect(AuxWorker& w) {
    std::vector<AuxJob> out;
    {
        std::lock_guard<std::mutex> lk(w.mu);
        out.swap(w.done);
    }
    w.pending -= int(out.size());
    return out;
}

// Registers an execution being prepared (busy, `est` bytes reserved).
Prepared* reserve(kz_server_s* s, kz_graph g, int kind, int k, uint64_t est) {
    Prepared* p = new Prepared();
    p->g = g;
    p->kind = kind;
    p->k = k;
    p->bytes = est;
    p->preparing = true;
    p->busy = 1;
    s->in_use += est;
    touch(s, p);
    s->prepared.push_back(p);
    return p;
}

// Turns a reservation into the execution's real size and marks it idle.
// Returns false (and releases it) when it then no longer fits next to the
// others.
bool settle(kz_server_s* s, Prepared* p, uint64_t bytes, double ms) {
    s->in_use -= p->bytes;
    p->bytes = bytes;
    s->in_use += bytes;
    p->cost_ms = ms;
    p->preparing = false;
    touch(s, p);
    const bool fits = make_room(s, 0);  // p is still busy: not a victim
    p->busy--;
    if (!fits) release_prepared(s, p);
    return fits;
}

// Takes the helper threads' finished jobs.  Queries finished here (failed
// preparations, census results) are appended to *gone.
void collect_aux(kz_server_s* s, std::vector<SQuery*>* gone) {
    for (AuxJob& j : aux_collect(s->prep)) {
        Prepared* p = j.p;
        const kz_graph g = p->g;
        DeviceGraph& d = g->g;
        if (j.rc && (j.rc == KZ_ERR_NOMEM || j.rc == KZ_ERR_DEVICE) && !p->retried && release_idle(s, p)) {
            // out of device memory despite the estimate: once more, with
            // every idle execution released
            p->retried = true;
            aux_submit(s->prep, j);
            continue;
        }
        if (j.rc) {
            p->preparing = false;
            p->busy--;
            release_prepared(s, p);
            if (!clique_prepared_on(s, g) && !find_prepared(s, g, kForwardCopy, 0) && d.fwd_ready &&
                !clique_exec_on(d))
                release_forward(d);
            finish(s, j.q, j.rc, j.err);
            gone->push_back(j.q);
            continue;
        }
        const uint64_t bytes =
            p->kind == kForwardCopy ? d.fwd_bytes : static_cast<CliqueJob*>(d.clique_exec[p->k].get())->bytes;
        if (!settle(s, p, bytes, j.ms) && !any_busy(s)) {
            finish(s, j.q, KZ_ERR_BUDGET);
            gone->push_back(j.q);
        }
    }
    for (AuxJob& j : aux_collect(s->census)) {
        Prepared* p = j.p;
        if (j.rc == KZ_ERR_DEVICE && !j.q->retried) {
            // a HIP runtime error (as preparations, once more; a lasting
            // one fails the second run too)
            j.q->retried = true;
            aux_submit(s->census, AuxJob{kJobCensus, j.q, p, 0, 0});
            continue;
        }
        if (j.rc) {
            p->busy--;
            p->preparing = false;
            if (!census_bytes(p->g->g)) release_prepared(s, p);
        } else if (p->preparing) {
            // the first census on this graph: its copy now has a real size
            settle(s, p, census_bytes(p->g->g), j.ms);
        } else {
            p->busy--;
            touch(s, p);
        }
        finish(s, j.q, j.rc, j.err);
    }
}

// One admission step for query q.  Returns 1 when q left the pending list
// (launched, handed to the census thread, or finished), 0 when it stays
// (being prepared, its execution busy, or no free stream: later queries may
// pass it), -1 when it must wait for memory (later queries wait too).
int admit(kz_server_s* s, SQuery* q, int* free_slot) {
    DeviceGraph& d = q->g->g;
    Prepared* p = find_prepared(s, q->g, q->kind, q->k);
    if (p && (p->busy || p->preparing)) return 0;
    auto no_room = [&]() {
        if (any_busy(s)) return -1;
        finish(s, q, KZ_ERR_BUDGET);  // nothing running would free more
        return 1;
    };
    if (q->kind == KZ_QUERY_CENSUS) {
        if (s->census.pending) return 0;
        if (!p) {
            const bool cached = bool(d.census);
            const uint64_t est = cached ? census_bytes(d) : estimate_bytes(d, q->kind, 0);
            if (est > budget_of(s)) {
                finish(s, q, KZ_ERR_BUDGET);
                return 1;
            }
            if (!make_room(s, est)) return no_room();
            p = reserve(s, q->g, q->kind, 0, est);
            // a copy kept from a direct call is adopted with its real size
            if (cached) p->preparing = false;
        } else {
            p->busy++;
        }
        aux_submit(s->census, AuxJob{kJobCensus, q, p, 0, 0});
        return 1;
    }
    if (!p) {
        Prepared* f = nullptr;
        if (!d.oriented) {
            f = find_prepared(s, q->g, kForwardCopy, 0);
            if (f && f->preparing) return 0;
            if (!f && !d.fwd_ready) {
                // a symmetric graph's first clique query: its degree
                // orientation first, then the job, sized from the exact
                // forward degrees
                const uint64_t est = forward_estimate(d);
                if (est > budget_of(s)) {
                    finish(s, q, KZ_ERR_BUDGET);
                    return 1;
                }
                if (!make_room(s, est)) return no_room();
                f = reserve(s, q->g, kForwardCopy, 0, est);
                aux_submit(s->prep, AuxJob{kJobForward, q, f, 0, 0});
                return 0;
            }
            if (!f && d.fwd_bytes) {  // built by a direct call: adopted
                f = reserve(s, q->g, kForwardCopy, 0, d.fwd_bytes);
                f->preparing = false;
                f->busy = 0;
            }
        }
        const bool cached = bool(d.clique_exec[q->k]);
        const uint64_t est = cached ? static_cast<CliqueJob*>(d.clique_exec[q->k].get())->bytes
                                    : estimate_bytes(d, q->kind, q->k);
        if (est > budget_of(s)) {
            finish(s, q, KZ_ERR_BUDGET);
            return 1;
        }
        if (f) f->busy++;  // the job needs it: not a victim for its room
        const bool room = make_room(s, est);
        if (f) f->busy--;
        if (!room) return no_room();
        p = reserve(s, q->g, q->kind, q->k, est);
        if (!cached) {
            aux_submit(s->prep, AuxJob{kJobPrepare, q, p, 0, 0});
            return 0;
        }
        p->preparing = false;
        p->busy = 0;
    }
    // prepared and idle: launch it on a free stream
    if (*free_slot < 0) return 0;
    CliqueJob* job = static_cast<CliqueJob*>(d.clique_exec[q->k].get());
    const int slot = *free_slot;
    p->busy++;
    touch(s, p);
    q->t_start = Clock::now();
    s->slot_query[slot] = q;
    s->slot_prep[slot] = p;
    if (!job || clique_job_launch(*job, s->streams[slot], s->pinned + slot) != KZ_OK ||
        hipEventRecord(s->events[slot], s->streams[slot]) != hipSuccess) {
        (void)hipGetLastError();
        p->busy--;
        s->slot_query[slot] = nullptr;
        s->slot_prep[slot] = nullptr;
        finish(s, q, KZ_ERR_DEVICE);
        return 1;
    }
    *free_slot = -1;
    for (int i = 0; i < int(s->slot_query.size()); ++i)
        if (!s->slot_query[i]) {
            *free_slot = i;
            break;
        }
    return 1;
}

}  // namespace
}  // namespace kz

using namespace kz;

static void worker_main(kz_server_s* s) {
    (void)hipSetDevice(s->device);
    const int nslots = int(s->streams.size());
    while (true) {
        // completions on the streams
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
        {
            std::lock_guard<std::mutex> lk(s->mu);
            s->aux_event = false;
        }
        // the helpers' finished jobs
        std::vector<SQuery*> gone;
        collect_aux(s, &gone);
        const bool helpers_idle = !s->prep.pending && !s->census.pending;
        // admission and launch, in arrival order
        std::vector<SQuery*> batch;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            if (s->pending.empty() && busy_slots == 0 && helpers_idle) {
                if (s->stopping) break;
                s->cv_work.wait(lk, [&] { return s->stopping || !s->pending.empty() || s->aux_event; });
                continue;
            }
            batch.assign(s->pending.begin(), s->pending.end());
        }
        int free_slot = -1;
        for (int i = 0; i < nslots; ++i)
            if (!s->slot_query[i]) {
                free_slot = i;
                break;
            }
        const size_t finished_by_helpers = gone.size();
        for (SQuery* q : batch) {
            bool left = false;
            for (size_t i = 0; i < finished_by_helpers; ++i) left |= gone[i] == q;
            if (left) continue;
            const int r = admit(s, q, &free_slot);
            if (r > 0) gone.push_back(q);
            if (r < 0) break;  // waiting for memory: later queries do not overtake it
        }
        if (!gone.empty()) {
            std::lock_guard<std::mutex> lk(s->mu);
            for (SQuery* q : gone)
                for (auto it = s->pending.begin(); it != s->pending.end(); ++it)
                    if (*it == q) {
                        s->pending.erase(it);
                        break;
                    }
        } else {
            std::unique_lock<std::mutex> lk(s->mu);
            if (!s->aux_event) s->cv_work.wait_for(lk, std::chrono::microseconds(50));
        }
    }
}

static void stop_helpers(kz_server_s* s) {
    for (AuxWorker* w : {&s->prep, &s->census}) {
        {
            std::lock_guard<std::mutex> lk(w->mu);
            w->stop = true;
        }
        w->cv.notify_all();
        if (w->th.joinable()) w->th.join();
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
    s->prep.th = std::thread(aux_loop, s, &s->prep);
    s->census.th = std::thread(aux_loop, s, &s->census);
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
    const std::string err = q->err;
    s->live.erase(ticket);
    delete q;
    if (out->status) set_error("%s", err.c_str());  // the failed query's message, for this thread
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
    stop_helpers(s);
    for (auto st : s->streams) (void)hipStreamDestroy(st);
    for (auto ev : s->events) (void)hipEventDestroy(ev);
    (void)hipHostFree(s->pinned);
    // the executions prepared under this server's budget are released,
    // clique executions before the forward copies they run on
    for (int pass = 0; pass < 2; ++pass)
        for (size_t i = s->prepared.size(); i-- > 0;)
            if ((s->prepared[i]->kind == kForwardCopy) == (pass == 1)) release_prepared(s, s->prepared[i]);
    for (auto& kv : s->live) delete kv.second;
    delete s;
    return KZ_OK;
}

}  // extern "C"
