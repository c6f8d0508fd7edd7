// DistributedApplication: the entry point the generated apps construct and
// run (include/distributed_application.h:38-65 of the reference, same
// constructor, run() and aggregator API).
//
// On this engine the constructor loads the .dgraph (kz_dgraph_load), selects
// the GPU (kz_init, device from $KZ_DEVICE, default 0) and keeps the whole
// CSR resident in HBM (kz_graph_upload).  run() probes the app's apply()
// callback into a plan and executes it on the GPU: clique plans with the
// persistent bitmap kernels, every other plan with the plain level-by-level
// executor.  Counts are added to the app's Aggregators exactly as the
// reference's threads would, so evaluate() and the multiplicity division in
// the app are unchanged.  $KZ_METHOD=plain forces the level-by-level form.
#ifndef DISTRIBUTED_APPLICATION_H
#define DISTRIBUTED_APPLICATION_H

#include <functional>
#include <string>
#include <vector>

#include "engine.h"
#include "graph.h"
#include "graph_loader.h"

struct kz_csr;  // include/kz.h

namespace Khuzdul {

typedef HashBasedNumaAwareDistributedCSRGraph DistGraph;

class DistributedApplication {
   private:
    int max_depth_;
    int max_pattern_size_;
    int num_threads_;
    int num_sockets_;
    int node_id_;
    int num_nodes_;
    DistGraph graph_;
    ApplyFunction apply_fun_;
    long long memory_size_per_level_;
    void* device_graph_;  // kz_graph
    void* paged_graph_;   // kz_paged: lists in host memory (beyond the resident budget)
    bool probe_only_;
    bool labelled_;  // load_vertex_labels: run() counts labelled patterns
    typedef void (*AddFn)(void*, uint64_t);
    std::vector<void*> aggregators_;  // in allocation order
    std::vector<AddFn> aggregator_add_;

    void register_aggregator(void* a, AddFn add);
    void set_host_labels(const kz_csr& csr);
    void run_labelled(bool restrict_starting_vertex_label, LabelId starting_vertex_label);
    void unregister_aggregator(void* a);

   public:
    DistributedApplication(int num_threads, std::string graph_path, int max_pattern_size, ApplyFunction apply_fun,
                           bool load_vertex_labels = false, bool finer_grained_multithreading = false,
                           long long memory_size_per_level = (long long)4 * 1024 * 1024 * 1024,
                           VertexId cache_degree_threshold = 64, double relative_cache_size = 0.15);
    virtual ~DistributedApplication();

    template <typename T>
    Aggregator<T>* alloc_aggregator() {
        Aggregator<T>* a = new Aggregator<T>(num_sockets_, num_threads_);
        register_aggregator(a, [](void* p, uint64_t d) { static_cast<Aggregator<T>*>(p)->add(0, 0, T(d)); });
        return a;
    }
    template <typename T>
    void dealloc_aggregator(Aggregator<T>* a) {
        unregister_aggregator(a);
        delete a;
    }
    void run(bool disable_output = false, bool restrict_starting_vertex_label = false, LabelId starting_vertex_label = 0);

    int get_num_threads() { return num_threads_; }
    int get_num_sockets() { return num_sockets_; }
    DistGraph* get_graph_obj(int socket_id = 0) {
        (void)socket_id;
        return &graph_;
    }
};

}  // namespace Khuzdul

#endif
