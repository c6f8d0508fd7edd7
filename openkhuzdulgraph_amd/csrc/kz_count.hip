// C-ABI entry points for pattern counting on resident graphs.
//
// kz_count_cliques replaces DistributedApplication::run
// (src/distributed_application.cc:259-356) + Aggregator<EdgeId>::evaluate
// (include/engine.h:628-645) for the clique apps: it returns the same uint64
// that the app prints after the multiplicity division (1 for cliques).
// kz_batch_* run many (graph, k) queries in one hipGraph launch; the
// reference runs one app and pattern per process.

#include <algorithm>
#include <chrono>
#include <cstring>
#include <memory>

#include "kz_device.h"
#include "kz_kernels.h"
#include "kz_plan.h"

namespace kz {

// For each vertex, the index of the first neighbour > v (symmetric graphs).
__global__ void fwd_start_kernel(const uint64_t* __restrict__ rowptr, const uint32_t* __restrict__ col, uint32_t n,
                                 uint32_t* __restrict__ skip, uint32_t* __restrict__ fdeg) {
    const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n) return;
    const uint64_t b = rowptr[v];
    const uint32_t d = uint32_t(rowptr[v + 1] - b);
    const uint32_t s = upper_bound_global(col + b, d, v);
    skip[v] = s;
    fdeg[v] = d - s;
}

// Wave per vertex copies the forward part of each list.
__global__ void fwd_copy_kernel(const uint64_t* __restrict__ rowptr, const uint32_t* __restrict__ col, uint32_t n,
                                const uint32_t* __restrict__ skip, const uint64_t* __restrict__ frowptr,
                                uint32_t* __restrict__ fcol) {
    const uint32_t v = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const uint32_t lane = threadIdx.x & 63;
    if (v >= n) return;
    const uint64_t src = rowptr[v] + skip[v];
    const uint64_t dst = frowptr[v];
    const uint32_t len = uint32_t(frowptr[v + 1] - dst);
    for (uint32_t i = lane; i < len; i += 64) fcol[dst + i] = col[src + i];
}

int ensure_forward(DeviceGraph& g) {
    if (g.fwd_ready) return KZ_OK;
    if (g.oriented) {
        g.fwd_rowptr = g.rowptr;
        g.fwd_col = g.col;
        g.h_fdeg = g.h_deg;
        g.fwd_bytes = 0;
        g.fwd_max = g.max_degree;
        g.fwd_ready = true;
        return KZ_OK;
    }
    // Counts do not depend on the vertex order; ranking by degree bounds
    // every forward degree by sqrt(2m) (hubs keep few forward neighbours).
    uint64_t fnnz = 0;
    if (int e = orient_resident(g, &g.fwd_rowptr, &g.fwd_col, nullptr, &g.h_fdeg, &fnnz)) return e;
    g.fwd_bytes = 8ull * (uint64_t(g.n) + 1) + col_alloc_bytes(fnnz);
    g.fwd_max = 0;
    for (uint32_t x : g.h_fdeg) g.fwd_max = std::max(g.fwd_max, x);
    g.fwd_ready = true;
    return KZ_OK;
}

void release_forward(DeviceGraph& g) {
    if (!g.fwd_ready || g.oriented) return;
    if (rt().ready) (void)hipStreamSynchronize(rt().stream);
    (void)hipFree(g.fwd_rowptr);
    (void)hipFree(g.fwd_col);
    g.fwd_rowptr = nullptr;
    g.fwd_col = nullptr;
    g.h_fdeg.clear();
    g.h_fdeg.shrink_to_fit();
    g.fwd_bytes = 0;
    g.fwd_max = 0;
    g.fwd_ready = false;
}

static BucketCfg graph_buckets(const DeviceGraph& d, int k) {
    if (!d.bucket_set[k]) return default_buckets();
    BucketCfg c;
    for (int b = 0; b < kNumBuckets; ++b) {
        c.hi[b] = d.bucket_hi[k][b];
        c.block[b] = d.bucket_block[k][b];
    }
    c.serial = d.bucket_serial[k];
    return c;
}

static void store_buckets(DeviceGraph& d, int k, const BucketCfg& c) {
    for (int b = 0; b < kNumBuckets; ++b) {
        d.bucket_hi[k][b] = c.hi[b];
        d.bucket_block[k][b] = c.block[b];
    }
    d.bucket_serial[k] = c.serial;
    d.bucket_set[k] = true;
    d.clique_exec[k].reset();
}

static BucketCfg from_abi(const kz_bucket_cfg& a) {
    BucketCfg c = default_buckets();
    for (int b = 0; b < 4; ++b) c.hi[b] = a.bound[b];
    for (int b = 0; b < 5; ++b) c.block[b] = a.block[b];
    c.serial = a.serial ? 1u : 0u;
    return c;
}

static void to_abi(const BucketCfg& c, kz_bucket_cfg* a) {
    for (int b = 0; b < 4; ++b) a->bound[b] = c.hi[b];
    for (int b = 0; b < 5; ++b) a->block[b] = c.block[b];
    a->serial = c.serial;
}

static std::shared_ptr<void> wrap_job(CliqueJob* j) {
    return std::shared_ptr<void>(j, [](void* p) { delete static_cast<CliqueJob*>(p); });
}

// Prepared single-graph job for (graph, k), built on first use.
int graph_job(DeviceGraph& d, int k, CliqueJob** out) {
    if (int e = ensure_forward(d)) return e;
    auto& slot = d.clique_exec[k];
    if (!slot) {
        CliqueJob* j = new (std::nothrow) CliqueJob();
        if (!j) KZ_FAIL(KZ_ERR_NOMEM, "clique job");
        slot = wrap_job(j);
        std::vector<CliqueQuery> q = {CliqueQuery{d.fwd_rowptr, d.fwd_col, &d.h_fdeg, k}};
        if (int e = clique_job_prepare(*j, q, graph_buckets(d, k))) {
            slot.reset();
            return e;
        }
    }
    *out = static_cast<CliqueJob*>(slot.get());
    return KZ_OK;
}

}  // namespace kz

using namespace kz;

struct kz_batch_s {
    std::vector<kz_graph> graphs;
    std::vector<int> ks;
    std::unique_ptr<CliqueJob> job;
    BucketCfg cfg = default_buckets();
};

extern "C" {

int kz_count_cliques(kz_graph g, int k, int method, uint64_t* count, kz_run_stats* stats) {
    if (!g || !count) KZ_FAIL(KZ_ERR_INVALID, "null argument");
    if (k < 3 || k > 5) KZ_FAIL(KZ_ERR_UNSUPPORTED, "clique size %d outside 3..5 (MAX_PATTERN_SIZE 5)", k);
    if (int e = ensure_ready()) return e;
    auto t0 = std::chrono::steady_clock::now();
    DeviceGraph& d = g->g;
    double dev_ms = 0;
    kz_run_stats st;
    std::memset(&st, 0, sizeof(st));
    if (method == KZ_METHOD_PERSISTENT) {
        if (int e = ensure_forward(d)) return e;
        // a root whose forward list cannot be staged in LDS even after the
        // orientation (a clique core of more than kMaxBitmapDegree vertices)
        // is counted by the plain form, which has no degree limit
        if (d.fwd_max > kMaxBitmapDegree) method = KZ_METHOD_PLAIN;
    }
    if (method == KZ_METHOD_PERSISTENT) {
        CliqueJob* j = nullptr;
        if (int e = graph_job(d, k, &j)) return e;
        if (int e = clique_job_run(*j, count, &dev_ms)) return e;
        st.launches = 1;
        st.tasks = j->ntasks;
    } else if (method == KZ_METHOD_PLAIN) {
        PlanDesc pd;
        plan_clique(k, &pd);
        uint64_t counts[KZ_MAX_AGGREGATORS] = {0};
        if (int e = plan_execute(d, pd, true, counts, &st)) return e;
        *count = counts[0];
        dev_ms = st.device_ms;
    } else {
        KZ_FAIL(KZ_ERR_INVALID, "unknown method %d", method);
    }
    auto t1 = std::chrono::steady_clock::now();
    st.device_ms = dev_ms;
    st.wall_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
    if (stats) *stats = st;
    return KZ_OK;
}

int kz_clique_get_buckets(kz_graph g, int k, kz_bucket_cfg* out) {
    if (!g || !out || k < 3 || k > 5) KZ_FAIL(KZ_ERR_INVALID, "bad argument");
    to_abi(graph_buckets(g->g, k), out);
    return KZ_OK;
}

int kz_clique_set_buckets(kz_graph g, int k, const kz_bucket_cfg* cfg) {
    if (!g || !cfg || k < 3 || k > 5) KZ_FAIL(KZ_ERR_INVALID, "bad argument");
    BucketCfg c = from_abi(*cfg);
    if (int e = check_buckets(c)) return e;
    store_buckets(g->g, k, c);
    return KZ_OK;
}

int kz_clique_tune(kz_graph g, int k, uint32_t reps, kz_bucket_cfg* chosen, double* best_ms) {
    if (!g || k < 3 || k > 5) KZ_FAIL(KZ_ERR_INVALID, "bad argument");
    if (int e = ensure_ready()) return e;
    DeviceGraph& d = g->g;
    if (reps == 0) reps = 1;
    double best = 1e300;
    BucketCfg best_cfg = default_buckets();
    for (const BucketCfg& c : tuning_candidates()) {
        store_buckets(d, k, c);
        CliqueJob* j = nullptr;
        if (int e = graph_job(d, k, &j)) return e;
        uint64_t cnt = 0;
        double ms = 0, tot = 0;
        if (int e = clique_job_run(*j, &cnt, &ms)) return e;  // warm-up
        for (uint32_t r = 0; r < reps; ++r) {
            if (int e = clique_job_run(*j, &cnt, &ms)) return e;
            tot += ms;
        }
        if (tot / reps < best) {
            best = tot / reps;
            best_cfg = c;
        }
    }
    store_buckets(d, k, best_cfg);
    if (chosen) to_abi(best_cfg, chosen);
    if (best_ms) *best_ms = best;
    return KZ_OK;
}

int kz_batch_create(const kz_graph* graphs, uint32_t num_graphs, const int* ks, uint32_t num_ks, kz_batch* out) {
    if (!graphs || !ks || !out || num_graphs == 0 || num_ks == 0) KZ_FAIL(KZ_ERR_INVALID, "bad argument");
    if (int e = ensure_ready()) return e;
    for (uint32_t j = 0; j < num_ks; ++j)
        if (ks[j] < 3 || ks[j] > 5) KZ_FAIL(KZ_ERR_UNSUPPORTED, "clique size %d outside 3..5", ks[j]);
    kz_batch b = new (std::nothrow) kz_batch_s();
    if (!b) KZ_FAIL(KZ_ERR_NOMEM, "batch");
    b->graphs.assign(graphs, graphs + num_graphs);
    b->ks.assign(ks, ks + num_ks);
    for (kz_graph g : b->graphs) {
        if (!g) {
            delete b;
            KZ_FAIL(KZ_ERR_INVALID, "null graph in batch");
        }
        if (int e = ensure_forward(g->g)) {
            delete b;
            return e;
        }
    }
    *out = b;
    return KZ_OK;
}

static int batch_prepare(kz_batch b) {
    if (b->job && b->job->ready) return KZ_OK;
    b->job.reset(new CliqueJob());
    std::vector<CliqueQuery> q;
    for (kz_graph g : b->graphs)
        for (int k : b->ks) q.push_back(CliqueQuery{g->g.fwd_rowptr, g->g.fwd_col, &g->g.h_fdeg, k});
    if (int e = clique_job_prepare(*b->job, q, b->cfg)) {
        b->job.reset();
        return e;
    }
    return KZ_OK;
}

int kz_batch_run(kz_batch b, uint64_t* counts, kz_run_stats* stats) {
    if (!b || !counts) KZ_FAIL(KZ_ERR_INVALID, "null argument");
    auto t0 = std::chrono::steady_clock::now();
    if (int e = batch_prepare(b)) return e;
    double ms = 0;
    if (int e = clique_job_run(*b->job, counts, &ms)) return e;
    if (stats) {
        std::memset(stats, 0, sizeof(*stats));
        stats->device_ms = ms;
        stats->launches = 1;
        stats->tasks = b->job->ntasks;
        stats->wall_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
    return KZ_OK;
}

int kz_batch_tune(kz_batch b, uint32_t reps, kz_bucket_cfg* chosen, double* best_ms) {
    if (!b) KZ_FAIL(KZ_ERR_INVALID, "null argument");
    if (reps == 0) reps = 1;
    std::vector<uint64_t> counts(b->graphs.size() * b->ks.size());
    double best = 1e300;
    BucketCfg best_cfg = default_buckets();
    for (const BucketCfg& c : tuning_candidates()) {
        b->cfg = c;
        b->job.reset();
        if (int e = batch_prepare(b)) return e;
        double ms = 0, tot = 0;
        if (int e = clique_job_run(*b->job, counts.data(), &ms)) return e;
        for (uint32_t r = 0; r < reps; ++r) {
            if (int e = clique_job_run(*b->job, counts.data(), &ms)) return e;
            tot += ms;
        }
        if (tot / reps < best) {
            best = tot / reps;
            best_cfg = c;
        }
    }
    b->cfg = best_cfg;
    b->job.reset();
    if (chosen) to_abi(best_cfg, chosen);
    if (best_ms) *best_ms = best;
    return KZ_OK;
}

int kz_batch_release(kz_batch b) {
    if (!b) return KZ_OK;
    if (rt().ready) (void)hipStreamSynchronize(rt().stream);
    delete b;
    return KZ_OK;
}

int kz_clique_workload(kz_graph g, uint64_t* out6) {
    if (!g || !out6) KZ_FAIL(KZ_ERR_INVALID, "null argument");
    if (int e = ensure_ready()) return e;
    DeviceGraph& d = g->g;
    if (int e = ensure_forward(d)) return e;
    CliqueView v;
    v.rowptr = d.fwd_rowptr;
    v.col = d.fwd_col;
    v.h_fdeg = d.h_fdeg;
    v.owner = &d;
    return clique_workload(v, out6);
}

}  // extern "C"
