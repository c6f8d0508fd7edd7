// Open and closed wedges at an edge with VertexSet::bounded_subtracted_by
// (src/vertex_set.cc:146-184 of the reference) and intersect_and_count
// (:88-111).  For every edge {v1 < v0} of a symmetric graph:
//   open   += |{v2 in N(v0) \ N(v1) : v2 < v1}|  (paths v1 - v0 - v2 with
//             v2 < v1 < v0 and v1, v2 not adjacent)
//   closed += |N(v0) & N(v1)|                    (three per triangle)
// so that open = sum_v C(d_<(v), 2) - triangles, with d_<(v) the number of
// neighbours below v.
//
// usage: example_bounded_wedges <graph.dgraph>
#include <cstdio>
#include <cstdlib>

#include "distributed_application.h"
#include "distributed_sys.h"
#include "engine.h"
#include "vertex_set.h"

namespace Khuzdul {

Aggregator<EdgeId>* open_wedges;
Aggregator<EdgeId>* closed_wedges;

void apply(ExtendableEmbedding& e, Context context, AbstractEmbeddingExplorationEngine* engine) {
    if (e.get_size() == 1) {
        const VertexId v0 = e.get_matched_vertex(0);
        VertexSet n0 = e.get_matched_vertex_nbrs(0);
        for (VertexId i = 0; i < n0.get_num_vertices(); ++i) {
            const VertexId v = n0.get_vertex(i);
            if (v < v0) engine->scatter(v, true, 0, e, context);
        }
        return;
    }
    const VertexId v1 = e.get_matched_vertex(1);
    VertexSet n0 = e.get_matched_vertex_nbrs(0);
    VertexSet n1 = e.get_matched_vertex_nbrs(1);
    VertexId* buf =
        (VertexId*)engine->alloc_thread_local_scratchpad_buffer(context, n0.get_num_vertices() * sizeof(VertexId));
    VertexSet open_set(buf, 0);
    n0.bounded_subtracted_by(&n1, &open_set, v1);
    open_wedges->add(context, open_set.get_num_vertices());
    closed_wedges->add(context, n0.intersect_and_count(&n1));
    engine->clear_thread_local_scratchpad_buffer(context);
}

}  // namespace Khuzdul

int main(int argc, char** argv) {
    if (argc < 2) {
        std::fprintf(stderr, "usage: %s <graph.dgraph>\n", argv[0]);
        return 1;
    }
    DistributedSys::init_distributed_sys();
    Khuzdul::DistributedApplication* app = new Khuzdul::DistributedApplication(1, argv[1], 3, Khuzdul::apply);
    Khuzdul::open_wedges = app->alloc_aggregator<EdgeId>();
    Khuzdul::closed_wedges = app->alloc_aggregator<EdgeId>();
    app->run();
    std::printf("Pattern count 0: %llu\n", (unsigned long long)Khuzdul::open_wedges->evaluate());
    std::printf("Pattern count 1: %llu\n", (unsigned long long)(Khuzdul::closed_wedges->evaluate() / 3));
    app->dealloc_aggregator(Khuzdul::open_wedges);
    app->dealloc_aggregator(Khuzdul::closed_wedges);
    delete app;
    DistributedSys::finalize_distributed_sys();
    return 0;
}
