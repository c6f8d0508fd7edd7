// Embedding-extension API seen by the generated apps (the API part of
// include/engine.h of the reference: lines 21-76, 582-673, 795-859,
// 1097-1263).  Same names, signatures and semantics:
//   * ApplyFunction (engine.h:589) is the app's per-embedding callback;
//   * cached objects are laid out [n_objs][offsets..][data] (engine.h:592-613);
//   * Aggregator<T> keeps one padded slot per (socket, thread) and evaluate()
//     sums them (engine.h:615-673);
//   * ExtendableEmbedding exposes matched vertices, their neighbour lists and
//     cached objects (engine.h:809-839);
//   * AbstractEmbeddingExplorationEngine is the callback's view of the
//     engine (engine.h:846-859).
// On this engine apply() is probed once per run into a plan that the MI355X
// kernels execute (distributed_application.cpp / probe.cpp).
#ifndef ENGINE_H
#define ENGINE_H

#include <assert.h>
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

#include <condition_variable>
#include <functional>
#include <mutex>
#include <set>
#include <thread>
#include <vector>

#include "distributed_sys.h"
#include "graph.h"
#include "graph_loader.h"
#include "kz_probe_hooks.h"
#include "shared_memory_sys.h"
#include "timer.h"
#include "utilities.h"

#define MAX_PATTERN_SIZE 5
#define MAX_DEPTH MAX_PATTERN_SIZE
#define MAX_NUM_INTERMEDIATE_OBJS 5
#define CHUNK_SIZE 64

typedef long long int lli;

struct Context {
    int socket_id;
    int thread_id;
    int depth;
    double* scatter_time;
} __attribute__((packed));

namespace Khuzdul {

class AbstractEmbeddingExplorationEngine;
class DistributedApplication;
struct ExtendableEmbedding;

typedef void (*ApplyFunction)(ExtendableEmbedding&, const Context, AbstractEmbeddingExplorationEngine*);

inline void init_cached_objs(VertexId* start_ptx) { *start_ptx = *(start_ptx + 1) = 0; }

inline VertexId* create_cached_obj(VertexId* start_ptx, VertexId size) {
    VertexId* data_ptx = start_ptx + MAX_NUM_INTERMEDIATE_OBJS + 2;
    VertexId num_objs = *start_ptx;
    assert(num_objs < MAX_NUM_INTERMEDIATE_OBJS);
    VertexId allocated_size = start_ptx[num_objs + 1];
    VertexId* new_obj_ptx = data_ptx + allocated_size;
    start_ptx[num_objs + 2] = allocated_size + size;
    (*start_ptx)++;
    if (__builtin_expect(kzk::g_hooks != nullptr, 0)) kzk::g_hooks->on_create_cached(start_ptx, new_obj_ptx, size);
    return new_obj_ptx;
}

inline VertexId* get_cached_obj(VertexId* start_ptx, int obj_id) {
    VertexId* data_ptx = start_ptx + MAX_NUM_INTERMEDIATE_OBJS + 2;
    VertexId num_objs = *start_ptx;
    assert(VertexId(obj_id) < num_objs);
    (void)num_objs;
    return data_ptx + start_ptx[obj_id + 1];
}

template <typename T>
class Aggregator {
   private:
    static const int padding = 128;  // bytes between slots (false-sharing guard)
    int num_threads_;
    int num_sockets_;
    T* local_reducers_[MAX_NUM_SOCKETS][MAX_NUM_THREADS];

   public:
    inline void clear() {
        for (int s = 0; s < num_sockets_; ++s)
            for (int t = 0; t < num_threads_; ++t) memset(local_reducers_[s][t], 0, sizeof(T));
    }
    inline T evaluate() {
        T r = 0;
        for (int s = 0; s < num_sockets_; ++s)
            for (int t = 0; t < num_threads_; ++t) r += *local_reducers_[s][t];
        T g = 0;
        MPI_Allreduce(&r, &g, 1, DistributedSys::get_mpi_data_type<T>(), MPI_SUM, MPI_COMM_WORLD);
        return g;
    }
    inline void add(int s_i, int t_i, T delta) { *local_reducers_[s_i][t_i] += delta; }
    inline void add(Context context, T delta) {
        if (__builtin_expect(kzk::g_hooks != nullptr, 0)) {
            kzk::g_hooks->on_add(this, uint64_t(delta));
            return;
        }
        add(context.socket_id, context.thread_id, delta);
    }
    Aggregator(int num_sockets, int num_threads) {
        num_sockets_ = num_sockets;
        num_threads_ = num_threads;
        for (int s = 0; s < num_sockets_; ++s)
            for (int t = 0; t < num_threads_; ++t)
                local_reducers_[s][t] = static_cast<T*>(aligned_alloc(64, sizeof(T) * (1 + padding / sizeof(T))));
        clear();
    }
    ~Aggregator() {
        for (int s = 0; s < num_sockets_; ++s)
            for (int t = 0; t < num_threads_; ++t) free(local_reducers_[s][t]);
    }
};

// Labelled patterns (include/engine.h:231-301 of the reference): the same
// fields and equality; the set of patterns keeps its interface (insert,
// is_exist, clear; the table size argument is accepted and not needed).
struct LabeledPattern {
    LabelId labels[MAX_PATTERN_SIZE];
    int size;
    bool operator==(const LabeledPattern& other) const {
        if (size != other.size) return false;
        for (int i = 0; i < size; ++i)
            if (labels[i] != other.labels[i]) return false;
        return true;
    }
} __attribute__((packed));

class LabeledPatternHashTable {
   private:
    std::set<std::vector<LabelId>> entries_;
    static std::vector<LabelId> key(const LabeledPattern& p) {
        std::vector<LabelId> k(1, LabelId(p.size));
        for (int i = 0; i < p.size; ++i) k.push_back(p.labels[i]);
        return k;
    }

   public:
    explicit LabeledPatternHashTable(size_t table_size = 1000000) { (void)table_size; }
    void clear() { entries_.clear(); }
    void insert(const LabeledPattern p) { entries_.insert(key(p)); }
    bool is_exist(const LabeledPattern p) { return entries_.count(key(p)) != 0; }
};

// Minimum-image domains of one labelled pattern (include/engine.h:405-580 of
// the reference: the same constructor, clear(), insert() and get_support()).
// A vertex joins domain i when apply() inserts it at pattern vertex i with
// enable set; the support is the smallest domain, at most the vertex count.
// On this engine apply() runs only while DistributedApplication::run()
// probes it: run() computes the domains on the GPU (kz_fsm_domains) and
// stores their sizes here.  An insert() outside a run (direct use of the
// class) goes to host bitmaps.
class DistributedPatternDomain {
   private:
    int max_pattern_size_;
    VertexId num_vertices_;
    bool from_engine_ = false;
    VertexId engine_sizes_[MAX_PATTERN_SIZE] = {};
    std::vector<uint64_t> bits_[MAX_PATTERN_SIZE];

   public:
    DistributedPatternDomain(int max_pattern_size, int num_threads, int num_sockets, VertexId num_vertices)
        : max_pattern_size_(max_pattern_size), num_vertices_(num_vertices) {
        (void)num_threads;
        (void)num_sockets;
        assert(max_pattern_size_ >= 1 && max_pattern_size_ <= MAX_PATTERN_SIZE);
    }
    void clear() {
        from_engine_ = false;
        for (int i = 0; i < MAX_PATTERN_SIZE; ++i) {
            engine_sizes_[i] = 0;
            bits_[i].clear();
        }
    }
    void insert(Context context, int pattern_vertex, VertexId input_graph_vertex, bool enable) {
        (void)context;
        if (__builtin_expect(kzk::g_hooks != nullptr, 0)) {
            kzk::g_hooks->on_domain_insert(this, pattern_vertex, input_graph_vertex, enable);
            return;
        }
        assert(pattern_vertex >= 0 && pattern_vertex < max_pattern_size_);
        if (!enable) return;
        std::vector<uint64_t>& b = bits_[pattern_vertex];
        if (b.empty()) b.assign((size_t(num_vertices_) + 63) / 64, 0);
        b[input_graph_vertex >> 6] |= 1ull << (input_graph_vertex & 63);
    }
    // the engine's domains for this run (run() calls it)
    void set_domain_sizes(const VertexId* sizes, int n) {
        clear();
        from_engine_ = true;
        for (int i = 0; i < n && i < MAX_PATTERN_SIZE; ++i) engine_sizes_[i] = sizes[i];
    }
    VertexId get_support(int pattern_size) {
        VertexId support = num_vertices_;
        for (int i = 0; i < pattern_size && i < MAX_PATTERN_SIZE; ++i) {
            VertexId d = 0;
            if (from_engine_) {
                d = engine_sizes_[i];
            } else {
                for (uint64_t w : bits_[i]) d += VertexId(__builtin_popcountll(w));
            }
            support = d < support ? d : support;
        }
        return support;
    }
};

struct ExtendableEmbedding {
    VertexId* matched_vertices_nbrs_ptx[MAX_PATTERN_SIZE];
    VertexId matched_vertices_num_nbrs[MAX_PATTERN_SIZE];
    VertexId matched_vertices[MAX_PATTERN_SIZE];
    int size;
    void* cached_objects[MAX_PATTERN_SIZE];
    void* compact_version;

    int get_size() { return size; }
    VertexId get_matched_vertex(int idx) {
        assert(idx >= 0 && idx < size);
        return matched_vertices[idx];
    }
    VertexSet get_matched_vertex_nbrs(int idx) {
        assert(idx >= 0 && idx < size);
        assert(matched_vertices_nbrs_ptx[idx] != nullptr && "Requested the graph data of a vertex marked as inactive!");
        return VertexSet(matched_vertices_nbrs_ptx[idx], matched_vertices_num_nbrs[idx]);
    }
    void* get_cached_obj(int dist) { return cached_objects[size - dist - 1]; }
};

class AbstractEmbeddingExplorationEngine {
   public:
    virtual void scatter(const VertexId new_v, const bool is_new_v_active, const uint32_t cached_obj_size,
                         ExtendableEmbedding& parent, const Context context, const bool is_finalized = false) = 0;
    virtual void* alloc_thread_local_scratchpad_buffer(const Context& context, const size_t& size) = 0;
    virtual void clear_thread_local_scratchpad_buffer(const Context& context) = 0;
    virtual LabelId get_vertex_label(const VertexId v) = 0;
    virtual ~AbstractEmbeddingExplorationEngine() {}
};

// Run metrics (include/engine.h:1117-1263).  On this engine they report the
// GPU runs: counting time, launches, embeddings per level, plan shape.
class LocalPerformanceMetric {
   public:
    static std::mutex mutex;
    static double load_graph_time;
    static double cpu_utilization;
    static double gpu_time_ms;
    static double probe_time_ms;
    static unsigned long long num_runs;
    static unsigned long long num_launches;
    static unsigned long long num_extensions;
    static std::string last_method;
    static void init_local_metrics();
    static void print_metrics(int runs = 1);
};

}  // namespace Khuzdul

#endif
