// Triangles counted per edge with VertexSet::intersect_and_count
// (src/vertex_set.cc:88-111 of the reference), an operation none of the
// generated apps calls.  Every edge {v1 < v0} of a symmetric graph adds
// |N(v0) & N(v1)|, the number of triangles on that edge, so the sum counts
// each triangle three times.
//
// usage: example_edge_triangles <graph.dgraph>
#include <cstdio>
#include <cstdlib>

#include "distributed_application.h"
#include "distributed_sys.h"
#include "engine.h"
#include "vertex_set.h"

namespace Khuzdul {

Aggregator<EdgeId>* edge_triangles;

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
    VertexSet n0 = e.get_matched_vertex_nbrs(0);
    VertexSet n1 = e.get_matched_vertex_nbrs(1);
    edge_triangles->add(context, n0.intersect_and_count(&n1));
}

}  // namespace Khuzdul

int main(int argc, char** argv) {
    if (argc < 2) {
        std::fprintf(stderr, "usage: %s <graph.dgraph>\n", argv[0]);
        return 1;
    }
    DistributedSys::init_distributed_sys();
    Khuzdul::DistributedApplication* app = new Khuzdul::DistributedApplication(1, argv[1], 3, Khuzdul::apply);
    Khuzdul::edge_triangles = app->alloc_aggregator<EdgeId>();
    app->run();
    const EdgeId sum = Khuzdul::edge_triangles->evaluate();
    std::printf("Pattern count 0: %llu\n", (unsigned long long)(sum / 3));
    app->dealloc_aggregator(Khuzdul::edge_triangles);
    delete app;
    DistributedSys::finalize_distributed_sys();
    return 0;
}
