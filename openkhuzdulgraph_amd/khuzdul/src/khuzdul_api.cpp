// libkhuzdul_amd.so: the Khuzdul application API implemented on the MI355X
// engine (libkz_hip.so).  The generated apps of the reference
// (distributed_applications/kautomine/*.cc, kgraphpi/*.cc) compile against
// the headers in ../include and link against this library unchanged.
//
// Replaces, per reference file:
//   src/distributed_application.cc:66-113  constructor: load the .dgraph,
//                                           bring the CSR into HBM
//   src/distributed_application.cc:259-356  run(): probe apply() into a plan
//                                           and count on the GPU
//   src/distributed_sys.cc                  single-process runtime (no MPI)
//   src/vertex_set.cc                       set operations (evaluated by the
//                                           probe; the GPU runs the real ones)
//   src/timer.cc:10-14                      get_time()
//   src/engine.cc:1809-1826                 LocalPerformanceMetric statics

#include <sys/stat.h>
#include <sys/time.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "distributed_application.h"
#include "distributed_sys.h"
#include "engine.h"
#include "kz.h"
#include "probe.h"
#include "timer.h"
#include "vertex_set.h"

// ----------------------------------------------------------------- runtime --

extern "C" {
int MPI_Barrier(MPI_Comm) { return MPI_SUCCESS; }
int MPI_Allreduce(const void* sendbuf, void* recvbuf, int count, MPI_Datatype type, MPI_Op, MPI_Comm) {
    size_t w = 4;
    if (type == MPI_CHAR) w = 1;
    else if (type == MPI_LONG_LONG || type == MPI_UNSIGNED_LONG || type == MPI_UNSIGNED_LONG_LONG || type == MPI_DOUBLE)
        w = 8;
    if (sendbuf != recvbuf) std::memcpy(recvbuf, sendbuf, w * size_t(count));
    return MPI_SUCCESS;
}
int MPI_Comm_rank(MPI_Comm, int* rank) {
    *rank = 0;
    return MPI_SUCCESS;
}
int MPI_Comm_size(MPI_Comm, int* size) {
    *size = 1;
    return MPI_SUCCESS;
}
}

DistributedSys* DistributedSys::instance_ = nullptr;
void DistributedSys::init_distributed_sys() {
    if (!instance_) instance_ = new DistributedSys();
}
void DistributedSys::finalize_distributed_sys() {
    delete instance_;
    instance_ = nullptr;
    kz_shutdown();
}
DistributedSys* DistributedSys::get_instance() {
    if (!instance_) init_distributed_sys();
    return instance_;
}

int SharedMemorySys::get_num_threads() { return 1; }

double get_time() {
    struct timeval tv;
    gettimeofday(&tv, nullptr);
    return double(tv.tv_sec) + double(tv.tv_usec) * 1e-6;
}

bool file_exits(const std::string& path) {
    struct stat st;
    return stat(path.c_str(), &st) == 0;
}
long file_size(const char* path) {
    struct stat st;
    return stat(path, &st) == 0 ? long(st.st_size) : -1;
}

void print_num_comparision() {}

// --------------------------------------------------------------- VertexSet --

void VertexSet::intersect_with(VertexSet* op, VertexSet* result) {
    kzk::probe_set_op(KZ_OP_INTERSECT, this, op, result);
}
void VertexSet::subtracted_by(VertexSet* op, VertexSet* result) { kzk::probe_set_op(KZ_OP_SUBTRACT, this, op, result); }
VertexId VertexSet::intersect_and_count(VertexSet* op) { return kzk::probe_count_op(this, op); }
void VertexSet::bounded_intersect_with(VertexSet* op, VertexSet* result, VertexId upper_bound) {
    kzk::probe_set_op(KZ_OP_INTERSECT, this, op, result, upper_bound);
}
void VertexSet::bounded_subtracted_by(VertexSet* op, VertexSet* result, VertexId upper_bound) {
    kzk::probe_set_op(KZ_OP_SUBTRACT, this, op, result, upper_bound);
}

// ------------------------------------------------------------------ metrics --

namespace Khuzdul {

std::mutex LocalPerformanceMetric::mutex;
double LocalPerformanceMetric::load_graph_time = 0;
double LocalPerformanceMetric::cpu_utilization = 0;
double LocalPerformanceMetric::gpu_time_ms = 0;
double LocalPerformanceMetric::probe_time_ms = 0;
unsigned long long LocalPerformanceMetric::num_runs = 0;
unsigned long long LocalPerformanceMetric::num_launches = 0;
unsigned long long LocalPerformanceMetric::num_extensions = 0;
std::string LocalPerformanceMetric::last_method;

void LocalPerformanceMetric::init_local_metrics() {
    gpu_time_ms = probe_time_ms = 0;
    num_runs = num_launches = num_extensions = 0;
}

void LocalPerformanceMetric::print_metrics(int runs) {
    if (runs <= 0) runs = 1;
    std::printf("\n**************** MI355X engine metrics ****************\n");
    std::printf("    graph load time:          %.3f (s)\n", load_graph_time);
    std::printf("    runs:                     %llu\n", num_runs);
    std::printf("    GPU counting time / run:  %.3f (ms)\n", gpu_time_ms / runs);
    std::printf("    plan probing time / run:  %.3f (ms)\n", probe_time_ms / runs);
    std::printf("    kernel or graph launches: %llu\n", num_launches);
    std::printf("    embeddings extended:      %llu\n", num_extensions);
    std::printf("    last execution form:      %s\n", last_method.c_str());
    std::printf("*******************************************************\n");
}

// ----------------------------------------------------- DistributedApplication --

static void die_kz(int rc, const char* what) {
    if (rc == KZ_OK) return;
    std::fprintf(stderr, "%s failed (%d): %s\n", what, rc, kz_last_error());
    std::exit(-1);
}

DistributedApplication::DistributedApplication(int num_threads, std::string graph_path, int max_pattern_size,
                                               ApplyFunction apply_fun, bool load_vertex_labels,
                                               bool finer_grained_multithreading, long long memory_size_per_level,
                                               VertexId cache_degree_threshold, double relative_cache_size) {
    (void)finer_grained_multithreading;
    assert(max_pattern_size >= 3);  // distributed_application.cc:71
    if (max_pattern_size > MAX_PATTERN_SIZE) {
        std::fprintf(stderr, "pattern size %d exceeds MAX_PATTERN_SIZE %d\n", max_pattern_size, MAX_PATTERN_SIZE);
        std::exit(-1);
    }
    labelled_ = load_vertex_labels;
    max_pattern_size_ = max_pattern_size;
    max_depth_ = max_pattern_size - 2;
    apply_fun_ = apply_fun;
    node_id_ = 0;
    num_nodes_ = 1;
    num_sockets_ = 1;
    // the thread count keeps the reference's meaning for Aggregator slots;
    // counting itself runs on the GPU
    num_threads_ = num_threads > 0 ? std::min(num_threads, MAX_NUM_THREADS) : 1;
    memory_size_per_level_ = memory_size_per_level;
    // KZ_PROBE_ONLY=1: extract and dump plans (KZ_PLAN_DUMP) without a GPU;
    // every count is then reported as 0.
    probe_only_ = std::getenv("KZ_PROBE_ONLY") != nullptr;
    device_graph_ = nullptr;
    paged_graph_ = nullptr;
    if (probe_only_) {
        std::printf("KZ_PROBE_ONLY: plans are extracted and dumped, nothing is counted\n");
        if (labelled_) {
            // the labelled probe needs the label count (the .dgraph is read
            // on the host; no GPU)
            kz_csr csr;
            std::memset(&csr, 0, sizeof(csr));
            die_kz(kz_dgraph_load(graph_path.c_str(), &csr), "kz_dgraph_load");
            graph_.num_vertices_ = csr.num_vertices;
            graph_.num_edges_ = csr.num_entries;
            graph_.max_degree_ = csr.max_degree;
            set_host_labels(csr);
            kz_csr_release(&csr);
        }
        return;
    }
    const char* dev = std::getenv("KZ_DEVICE");
    die_kz(kz_init(dev ? std::atoi(dev) : 0), "kz_init");
    if (memory_size_per_level > 0) die_kz(kz_set_level_budget(uint64_t(memory_size_per_level)), "kz_set_level_budget");

    std::printf("Loading graph %s...\n", graph_path.c_str());
    double t = -get_time();
    kz_csr csr;
    std::memset(&csr, 0, sizeof(csr));
    die_kz(kz_dgraph_load(graph_path.c_str(), &csr), "kz_dgraph_load");
    std::printf("    Number of vertices: %u\n", csr.num_vertices);
    std::printf("    Number of edges: %lu\n", (unsigned long)csr.num_entries);
    std::printf("    Max degree: %u\n", csr.max_degree);
    graph_.num_vertices_ = csr.num_vertices;
    graph_.num_edges_ = csr.num_entries;
    graph_.max_degree_ = csr.max_degree;
    graph_.oriented_ = csr.oriented != 0;
    if (labelled_) {
        set_host_labels(csr);
        std::printf("    Number of labels: %u\n", graph_.num_labels_);
    }
    // KZ_RESIDENT_BYTES: HBM budget for the neighbour lists.  A graph whose
    // lists exceed it keeps them in host memory and is counted in root
    // batches (kz_paged_*), with cache_degree_threshold and
    // relative_cache_size sizing the HBM cache of high-degree lists as the
    // reference's remote-data cache (src/engine.cc:94-510).
    const char* rb = std::getenv("KZ_RESIDENT_BYTES");
    const uint64_t resident = rb ? std::strtoull(rb, nullptr, 10) : 0;
    if (resident && 4 * csr.num_entries > resident) {
        if (labelled_) {
            std::fprintf(stderr, "a labelled graph must fit KZ_RESIDENT_BYTES\n");
            std::exit(-1);
        }
        kz_csr ori;
        std::memset(&ori, 0, sizeof(ori));
        const kz_csr* src = &csr;
        if (!csr.oriented) {
            die_kz(kz_orient_by_degree(&csr, &ori), "kz_orient_by_degree");
            src = &ori;
        }
        kz_paged_cfg cfg;
        std::memset(&cfg, 0, sizeof(cfg));
        cfg.list_bytes = resident;
        cfg.cache_bytes = uint64_t(std::max(0.0, relative_cache_size) * 4.0 * double(src->num_entries));
        cfg.cache_degree_threshold = cache_degree_threshold;
        kz_paged pg = nullptr;
        die_kz(kz_paged_create(src, &cfg, &pg), "kz_paged_create");
        kz_csr_release(&ori);
        kz_csr_release(&csr);
        paged_graph_ = pg;
        t += get_time();
        LocalPerformanceMetric::load_graph_time = t;
        std::printf("The neighbour lists exceed KZ_RESIDENT_BYTES (%llu): they stay in host memory, %llu bytes of "
                    "HBM hold one root batch's lists and %llu the cache (%.3f s to load).\n",
                    (unsigned long long)resident, (unsigned long long)cfg.list_bytes,
                    (unsigned long long)cfg.cache_bytes, t);
        return;
    }
    kz_graph g = nullptr;
    die_kz(kz_graph_upload(&csr, &g), "kz_graph_upload");
    kz_csr_release(&csr);
    device_graph_ = g;
    t += get_time();
    LocalPerformanceMetric::load_graph_time = t;
    std::printf("It takes %.3f seconds to load the graph into HBM.\n", t);
}

// Labels of a .dgraph read with load_vertex_labels: the label count and
// each label's vertices (graph.h:108-114 of the reference).  The device
// copy is attached by kz_graph_upload (a kz_csr carrying labels).
void DistributedApplication::set_host_labels(const kz_csr& csr) {
    if (!csr.has_labels || (!csr.labels && csr.num_vertices)) {
        std::fprintf(stderr, "load_vertex_labels: the graph file carries no vertex labels\n");
        std::exit(-1);
    }
    LabelId L = 0;
    for (VertexId v = 0; v < csr.num_vertices; ++v) L = std::max<LabelId>(L, csr.labels[v] + 1);
    graph_.num_labels_ = L;
    graph_.labeled_vertices_.assign(L, std::vector<VertexId>());
    for (VertexId v = 0; v < csr.num_vertices; ++v) graph_.labeled_vertices_[csr.labels[v]].push_back(v);
}

// One run() of a labelled app: apply() is probed into a labelled pattern
// (probe.cpp), its minimum-image domains come from the GPU
// (kz_fsm_domains, on the kz_fsm_mine level of the pattern's shape) and go
// into the app's DistributedPatternDomain, whose get_support() the app then
// reads (three_fsm.cc:572-587).  Shapes: edge, wedge, triangle, four-chain,
// three-star (every pattern kautomine/three_fsm.cc mines).
void DistributedApplication::run_labelled(bool restrict_starting_vertex_label, LabelId starting_vertex_label) {
    kzk::LabeledPlan lp;
    std::string err;
    double t = -get_time();
    if (kzk::extract_labeled_pattern(apply_fun_, max_pattern_size_, graph_.num_labels_, restrict_starting_vertex_label,
                                     starting_vertex_label, &lp, &err) != 0) {
        std::fprintf(stderr, "cannot turn the labelled apply() into a GPU pattern: %s\n", err.c_str());
        std::exit(-1);
    }
    t += get_time();
    LocalPerformanceMetric::probe_time_ms += t * 1e3;
    if (lp.empty || !lp.domain) return;  // no embedding: the app's cleared domain gives support 0
    // the kz_fsm_mine level and its position order (the KZ_FSM_* comments)
    const int k = lp.size;
    int deg[MAX_PATTERN_SIZE] = {0};
    int edges = 0;
    for (int i = 0; i < k; ++i) {
        deg[i] = __builtin_popcount(lp.adj[i]);
        edges += deg[i];
    }
    edges /= 2;
    int level = -1;
    int order[4] = {0, 1, 2, 3};  // level position -> pattern vertex
    auto other = [&](int i, int not1, int not2) {
        for (int j = 0; j < k; ++j)
            if ((lp.adj[i] >> j & 1u) && j != not1 && j != not2) return j;
        return -1;
    };
    if (k == 2 && edges == 1) {
        level = KZ_FSM_EDGE;
    } else if (k == 3 && edges == 3) {
        level = KZ_FSM_TRIANGLE;
    } else if (k == 3 && edges == 2) {
        level = KZ_FSM_WEDGE;  // center, leaf, leaf
        const int c = deg[0] == 2 ? 0 : (deg[1] == 2 ? 1 : 2);
        order[0] = c;
        order[1] = other(c, -1, -1);
        order[2] = other(c, order[1], -1);
    } else if (k == 4 && edges == 3) {
        int c = -1, leaves = 0;
        for (int i = 0; i < 4; ++i) {
            if (deg[i] == 3) c = i;
            leaves += deg[i] == 1;
        }
        if (c >= 0) {
            level = KZ_FSM_THREE_STAR;  // leaf, center, leaf, leaf
            order[1] = c;
            int o = 0;
            for (int i = 0; i < 4; ++i)
                if (i != c) order[o == 0 ? 0 : o + 1] = i, ++o;
        } else if (leaves == 2) {
            level = KZ_FSM_FOUR_CHAIN;  // path order[2] - order[0] - order[1] - order[3]
            int p = -1;
            for (int i = 0; i < 4 && p < 0; ++i)
                if (deg[i] == 2) p = i;
            const int n1 = other(p, -1, -1), n2 = other(p, n1, -1);
            const int q = deg[n1] == 2 ? n1 : n2;  // p's inner neighbour
            order[0] = p;
            order[1] = q;
            order[2] = q == n1 ? n2 : n1;
            order[3] = other(q, p, -1);
        }
    }
    if (level < 0) {
        std::fprintf(stderr, "labelled pattern of %d vertices and %d edges: only edges, wedges, triangles, "
                             "four-chains and three-stars are counted on the GPU\n", k, edges);
        std::exit(-1);
    }
    uint32_t labels[4] = {0, 0, 0, 0}, sizes[4] = {0, 0, 0, 0};
    for (int i = 0; i < k; ++i) labels[i] = lp.labels[order[i]];
    if (const char* tr = std::getenv("KZ_LABELLED_TRACE")) {
        // one line per run: the recovered pattern and its kz_fsm_mine level
        if (FILE* f = std::fopen(tr, "a")) {
            std::fprintf(f, "level %d labels", level);
            for (int i = 0; i < k; ++i) std::fprintf(f, " %u", labels[i]);
            std::fprintf(f, " adj");
            for (int i = 0; i < k; ++i) std::fprintf(f, " %u", lp.adj[i]);
            std::fprintf(f, "\n");
            std::fclose(f);
        }
    }
    if (probe_only_) {
        // nothing is counted: every pattern is reported with the largest
        // support (the vertex count), so that every level's candidates are
        // probed
        VertexId dom[MAX_PATTERN_SIZE];
        for (int i = 0; i < MAX_PATTERN_SIZE; ++i) dom[i] = graph_.num_vertices_;
        static_cast<Khuzdul::DistributedPatternDomain*>(lp.domain)->set_domain_sizes(dom, k);
        return;
    }
    kz_graph g = static_cast<kz_graph>(device_graph_);
    double tg = -get_time();
    die_kz(kz_fsm_domains(g, uint32_t(level), labels, sizes), "kz_fsm_domains");
    tg += get_time();
    VertexId dom[MAX_PATTERN_SIZE] = {0};
    for (int i = 0; i < k; ++i) dom[order[i]] = sizes[i];
    static_cast<Khuzdul::DistributedPatternDomain*>(lp.domain)->set_domain_sizes(dom, k);
    LocalPerformanceMetric::gpu_time_ms += tg * 1e3;
    LocalPerformanceMetric::num_runs++;
    LocalPerformanceMetric::num_launches++;
    LocalPerformanceMetric::last_method = "labelled pattern domains (kz_fsm_domains)";
}

DistributedApplication::~DistributedApplication() {
    if (device_graph_) kz_graph_release(static_cast<kz_graph>(device_graph_));
    if (paged_graph_) kz_paged_release(static_cast<kz_paged>(paged_graph_));
    device_graph_ = nullptr;
    paged_graph_ = nullptr;
}

void DistributedApplication::register_aggregator(void* a, AddFn add) {
    aggregators_.push_back(a);
    aggregator_add_.push_back(add);
}

void DistributedApplication::unregister_aggregator(void* a) {
    for (size_t i = 0; i < aggregators_.size(); ++i)
        if (aggregators_[i] == a) {
            aggregators_.erase(aggregators_.begin() + long(i));
            aggregator_add_.erase(aggregator_add_.begin() + long(i));
            return;
        }
}

static bool same_plan(const kz_plan_desc& a, const kz_plan_desc& b) {
    if (a.pattern_size != b.pattern_size || a.num_step_lists != b.num_step_lists) return false;
    if (std::memcmp(a.type_table, b.type_table, sizeof(a.type_table)) != 0) return false;
    for (uint32_t i = 0; i < a.num_step_lists; ++i) {
        const kz_step_list& x = a.lists[i];
        const kz_step_list& y = b.lists[i];
        if (x.num_ops != y.num_ops || x.num_outs != y.num_outs) return false;
        if (std::memcmp(x.ops, y.ops, sizeof(kz_op) * x.num_ops) != 0) return false;
        for (int o = 0; o < x.num_outs; ++o) {
            const kz_out& p = x.outs[o];
            const kz_out& q = y.outs[o];
            if (p.kind != q.kind || p.pred != q.pred || std::memcmp(&p.src, &q.src, sizeof(p.src)) != 0) return false;
            if (p.kind == KZ_OUT_COUNT && p.aggregator != q.aggregator) return false;
        }
    }
    return true;
}

void DistributedApplication::run(bool disable_output, bool restrict_starting_vertex_label,
                                 LabelId starting_vertex_label) {
    if (labelled_) {
        run_labelled(restrict_starting_vertex_label, starting_vertex_label);
        return;
    }
    if (restrict_starting_vertex_label) {
        std::fprintf(stderr, "label-restricted runs need the graph's labels (load_vertex_labels)\n");
        std::exit(-1);
    }
    kz_plan_desc plan;
    std::string err;
    double t = -get_time();
    if (kzk::extract_plan(apply_fun_, max_pattern_size_, aggregators_, &plan, &err) != 0) {
        std::fprintf(stderr, "cannot turn apply() into a GPU plan: %s\n", err.c_str());
        std::exit(-1);
    }
    t += get_time();
    LocalPerformanceMetric::probe_time_ms += t * 1e3;
    die_kz(kz_plan_validate(&plan), "kz_plan_validate");
    if (const char* dump = std::getenv("KZ_PLAN_DUMP")) {
        // appends the raw kz_plan_desc of this run (one record per run())
        if (FILE* f = std::fopen(dump, "ab")) {
            std::fwrite(&plan, sizeof(plan), 1, f);
            std::fclose(f);
        }
    }
    if (probe_only_) return;

    kz_graph g = static_cast<kz_graph>(device_graph_);
    uint64_t counts[KZ_MAX_AGGREGATORS] = {0};
    kz_run_stats st;
    std::memset(&st, 0, sizeof(st));
    const char* method = std::getenv("KZ_METHOD");
    const bool force_plain = method && std::strcmp(method, "plain") == 0;
    kz_plan_desc clique;
    bool is_clique = plan.num_aggregators == 1 && max_pattern_size_ <= 5 &&
                     kz_plan_clique(max_pattern_size_, &clique) == KZ_OK && same_plan(plan, clique);
    if (paged_graph_) {
        if (!is_clique) {
            std::fprintf(stderr, "a graph beyond KZ_RESIDENT_BYTES is counted by the clique apps only\n");
            std::exit(-1);
        }
        uint64_t c = 0;
        kz_paged_stats ps;
        die_kz(kz_paged_count_cliques(static_cast<kz_paged>(paged_graph_), max_pattern_size_, &c, &ps),
               "kz_paged_count_cliques");
        counts[0] = c;
        st.device_ms = ps.device_ms;
        st.launches = ps.batches;
        LocalPerformanceMetric::last_method = "paged clique kernels (lists fetched from host memory per root batch)";
    } else if (is_clique && !force_plain) {
        uint64_t c = 0;
        die_kz(kz_count_cliques(g, max_pattern_size_, KZ_METHOD_PERSISTENT, &c, &st), "kz_count_cliques");
        counts[0] = c;
        LocalPerformanceMetric::last_method = "persistent clique kernels (one hipGraph launch)";
    } else {
        die_kz(kz_count_plan(g, &plan, counts, KZ_MAX_AGGREGATORS, &st), "kz_count_plan");
        LocalPerformanceMetric::last_method = "plain level-by-level plan executor";
    }
    for (size_t i = 0; i < aggregators_.size() && i < KZ_MAX_AGGREGATORS; ++i)
        aggregator_add_[i](aggregators_[i], counts[i]);
    LocalPerformanceMetric::gpu_time_ms += st.device_ms;
    LocalPerformanceMetric::num_launches += st.launches;
    for (int s = 0; s <= KZ_MAX_PATTERN; ++s) LocalPerformanceMetric::num_extensions += st.embeddings[s];
    LocalPerformanceMetric::num_runs += 1;
    if (!disable_output)
        std::printf("node: 0, GPU runtime: %.3f (ms), %s, %u step lists\n", st.device_ms,
                    LocalPerformanceMetric::last_method.c_str(), plan.num_step_lists);
}

}  // namespace Khuzdul
