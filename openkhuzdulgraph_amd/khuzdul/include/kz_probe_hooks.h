// Observation points used while DistributedApplication::run() turns an
// app's apply() callback into a GPU plan (see probe.cpp).  Outside that
// probing pass the pointer is null and every hook is a single predicted
// branch; apply() itself never runs on the counting path.
#ifndef KZ_PROBE_HOOKS_H
#define KZ_PROBE_HOOKS_H

#include <stdint.h>

#include "types.h"

namespace kzk {

struct ProbeHooks {
    virtual void on_read(const VertexId* list, VertexId n) = 0;
    virtual void on_trim(const VertexId* list, VertexId n) = 0;
    virtual void on_add(const void* aggregator, uint64_t delta) = 0;
    virtual void on_create_cached(VertexId* start, VertexId* obj, VertexId size) = 0;
    // DistributedPatternDomain::insert from apply() (labelled mining)
    virtual void on_domain_insert(const void* domain, int pattern_vertex, VertexId v, bool enable) {
        (void)domain;
        (void)pattern_vertex;
        (void)v;
        (void)enable;
    }
    // VertexSet::get_num_vertices: a read when the list is one the probe
    // placed (a count added as a set's size), ignored otherwise
    virtual void on_size(const VertexId* list, VertexId n) {
        (void)list;
        (void)n;
    }
    virtual ~ProbeHooks() {}
};

extern ProbeHooks* g_hooks;

}  // namespace kzk

#endif
