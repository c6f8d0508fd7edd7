// 4-cliques with VertexSet::bounded_intersect_with (src/vertex_set.cc:113-144
// of the reference) doing the symmetry breaking: vertices are matched in
// decreasing order, each candidate set is cut below the vertex matched last.
// The size-2 embedding caches N(v0) & N(v1) below v1 as its prefix (the
// cached-object layout of the kautomine apps); size 3 intersects that prefix
// with N(v2) below v2 and adds the result's size.
//
// usage: example_bounded_four_clique <graph.dgraph>
#include <cstdio>
#include <cstdlib>

#include "distributed_application.h"
#include "distributed_sys.h"
#include "engine.h"
#include "vertex_set.h"

namespace Khuzdul {

Aggregator<EdgeId>* cliques;

void apply(ExtendableEmbedding& e, Context context, AbstractEmbeddingExplorationEngine* engine) {
    switch (e.get_size()) {
        case 1: {
            const VertexId v0 = e.get_matched_vertex(0);
            VertexSet n0 = e.get_matched_vertex_nbrs(0);
            const VertexId prefix_bytes =
                (n0.get_num_vertices() + 1 + MAX_NUM_INTERMEDIATE_OBJS + 2) * VertexId(sizeof(VertexId));
            for (VertexId i = 0; i < n0.get_num_vertices(); ++i) {
                const VertexId v = n0.get_vertex(i);
                if (v < v0) engine->scatter(v, true, prefix_bytes, e, context);
            }
            break;
        }
        case 2: {
            const VertexId v1 = e.get_matched_vertex(1);
            VertexSet n0 = e.get_matched_vertex_nbrs(0);
            VertexSet n1 = e.get_matched_vertex_nbrs(1);
            VertexId* objs = (VertexId*)e.get_cached_obj(0);
            init_cached_objs(objs);
            VertexId* list = create_cached_obj(objs, n0.get_num_vertices()) + 1;
            VertexSet below_v1(list, 0);
            n0.bounded_intersect_with(&n1, &below_v1, v1);
            *(list - 1) = below_v1.get_num_vertices();
            for (VertexId i = 0; i < below_v1.get_num_vertices(); ++i)
                engine->scatter(below_v1.get_vertex(i), true, 0, e, context);
            break;
        }
        case 3: {
            const VertexId v2 = e.get_matched_vertex(2);
            VertexId* cached = get_cached_obj((VertexId*)e.get_cached_obj(1), 0);
            VertexSet below_v1(cached + 1, *cached);
            VertexSet n2 = e.get_matched_vertex_nbrs(2);
            VertexId* buf = (VertexId*)engine->alloc_thread_local_scratchpad_buffer(
                context, below_v1.get_num_vertices() * sizeof(VertexId));
            VertexSet below_v2(buf, 0);
            below_v1.bounded_intersect_with(&n2, &below_v2, v2);
            cliques->add(context, below_v2.get_num_vertices());
            engine->clear_thread_local_scratchpad_buffer(context);
            break;
        }
        default:
            std::fprintf(stderr, "invalid embedding size %d\n", e.get_size());
            std::exit(-1);
    }
}

}  // namespace Khuzdul

int main(int argc, char** argv) {
    if (argc < 2) {
        std::fprintf(stderr, "usage: %s <graph.dgraph>\n", argv[0]);
        return 1;
    }
    DistributedSys::init_distributed_sys();
    Khuzdul::DistributedApplication* app = new Khuzdul::DistributedApplication(1, argv[1], 4, Khuzdul::apply);
    Khuzdul::cliques = app->alloc_aggregator<EdgeId>();
    app->run();
    std::printf("Pattern count 0: %llu\n", (unsigned long long)Khuzdul::cliques->evaluate());
    app->dealloc_aggregator(Khuzdul::cliques);
    delete app;
    DistributedSys::finalize_distributed_sys();
    return 0;
}
