// .dgraph format description (include/graph_loader.h:31-42 of the
// reference).  Loading is done by kz_dgraph_load in libkz_hip.so.
#ifndef GRAPH_LOADER_H
#define GRAPH_LOADER_H

#include "graph.h"
#include "types.h"

#define MAX_NUM_LABELS 48

struct GraphMetaDataStructOnDisk {
    VertexId num_vertices;
    EdgeId num_edges;
    int has_node_label;
    int has_edge_label;
} __attribute__((packed));

class JointHashBasedNumaAwareDistGraphLoader {};

#endif
