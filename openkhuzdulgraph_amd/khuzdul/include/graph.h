// Graph handle seen by apps (include/graph.h of the reference).  On this
// engine the whole CSR lives in HBM (libkz_hip.so kz_graph); the host keeps
// only its shape.  HUB_VERTEX_DEGREE_TH is kept for source compatibility.
#ifndef GRAPH_H
#define GRAPH_H

#include <stdint.h>

#include <vector>

#include "types.h"
#include "vertex_set.h"

#define HUB_VERTEX_DEGREE_TH (1024 * 1024 * 1024)

struct Empty {
    operator VertexId() const { return 0; }
} __attribute__((packed));

class HashBasedNumaAwareDistributedCSRGraph {
   public:
    VertexId num_vertices_ = 0;
    EdgeId num_edges_ = 0;
    VertexId max_degree_ = 0;
    bool oriented_ = false;
    // labelled graphs (load_vertex_labels): labels 0 .. num_labels_ - 1 and
    // the vertices of each label, ascending (graph.h:108-114, 409-421)
    LabelId num_labels_ = 0;
    std::vector<std::vector<VertexId>> labeled_vertices_;
    inline VertexId get_num_vertices() { return num_vertices_; }
    inline EdgeId get_num_edges() { return num_edges_; }
    inline VertexId get_max_degree() { return max_degree_; }
    inline LabelId get_num_labels() const { return num_labels_; }
    inline VertexSet get_labeled_vertices_set(const LabelId label) {
        if (label >= labeled_vertices_.size()) return VertexSet(nullptr, 0);
        std::vector<VertexId>& v = labeled_vertices_[label];
        return VertexSet(v.data(), VertexId(v.size()));
    }
};

#endif
