// VertexSet: a non-owning sorted list of vertex ids (include/vertex_set.h:14-55
// of the reference, same members and methods).
//
// intersect_with / subtracted_by keep the reference's contract
// (src/vertex_set.cc:20-86: sorted A & B, sorted A \ B, written to the
// result's list, length stored in the result), and so do their bounded
// forms (src/vertex_set.cc:113-184: only ids below upper_bound, which must
// be a matched vertex here) and intersect_and_count (:88-111, |A & B|).
// On this engine they are executed by the GPU kernels of libkz_hip.so as
// part of a plan; the host-side symbols exist so that an unchanged apply()
// links, and are evaluated only while DistributedApplication::run() probes
// apply() into a plan.  Calling them from host code outside apply() is
// reported as an error.  get_num_vertices() reports a read of the list, so
// a count added as a set's size is attributed to that set.
#ifndef VERTEX_SET_H
#define VERTEX_SET_H

#include <functional>
#include <iostream>
#include <string>

#include "dwarves_debug.h"
#include "kz_probe_hooks.h"
#include "memory_pool.h"
#include "types.h"

struct VertexSet {
    VertexId* vertices_list_;
    VertexId num_vertices_;

    VertexSet() {}
    VertexSet(VertexId* vertices_list, VertexId num_vertices)
        : vertices_list_(vertices_list), num_vertices_(num_vertices) {}

    inline VertexId* get_vertices_list() {
        if (__builtin_expect(kzk::g_hooks != nullptr, 0)) kzk::g_hooks->on_read(vertices_list_, num_vertices_);
        return vertices_list_;
    }
    inline VertexId get_num_vertices() const {
        if (__builtin_expect(kzk::g_hooks != nullptr, 0)) kzk::g_hooks->on_size(vertices_list_, num_vertices_);
        return num_vertices_;
    }
    inline VertexId get_vertex(VertexId idx) {
        if (__builtin_expect(kzk::g_hooks != nullptr, 0)) kzk::g_hooks->on_read(vertices_list_, num_vertices_);
        return vertices_list_[idx];
    }
    inline void set_vertices_list(VertexId* vertices_list) { vertices_list_ = vertices_list; }
    inline void set_num_vertices(VertexId num_vertices) { num_vertices_ = num_vertices; }
    inline void set_vertex(VertexId idx, VertexId vtx) { vertices_list_[idx] = vtx; }

    void intersect_with(VertexSet* op, VertexSet* result);
    void subtracted_by(VertexSet* op, VertexSet* result);
    VertexId intersect_and_count(VertexSet* op);
    void bounded_intersect_with(VertexSet* op, VertexSet* result, VertexId upper_bound);
    void bounded_subtracted_by(VertexSet* op, VertexSet* result, VertexId upper_bound);

    friend std::ostream& operator<<(std::ostream& os, const VertexSet& vs) {
        os << "vertex set: (size: " << vs.num_vertices_ << ") ";
        for (VertexId i = 0; i < vs.num_vertices_; ++i) os << vs.vertices_list_[i] << " ";
        return os;
    }
};

void print_num_comparision();

#endif
