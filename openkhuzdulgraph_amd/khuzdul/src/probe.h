// Plan extraction from an apply() callback (see probe.cpp).
#ifndef KZ_KHUZDUL_PROBE_H
#define KZ_KHUZDUL_PROBE_H

#include <string>
#include <vector>

#include "engine.h"
#include "kz.h"
#include "vertex_set.h"

namespace kzk {

// 0 on success; -1 with *error set when apply() does something the plan IR
// cannot express.
int extract_plan(Khuzdul::ApplyFunction fn, int max_pattern_size, const std::vector<void*>& aggregators,
                 kz_plan_desc* plan, std::string* error);

// A labelled pattern recovered from a labelled app's apply(): `size`
// vertices, adj[i] the bitmask of i's neighbours, labels[i] its label, and
// the DistributedPatternDomain apply() inserts into.  `empty`: no embedding
// can exist (a pattern vertex accepts no label, or the root label is not
// the pattern's).
struct LabeledPlan {
    int size = 0;
    uint32_t adj[MAX_PATTERN_SIZE] = {};
    LabelId labels[MAX_PATTERN_SIZE] = {};
    void* domain = nullptr;
    bool empty = false;
};

// 0 on success; -1 with *error set when apply() is not a labelled pattern
// walk the probe can recover.  root_known: the run restricts the roots to
// root_label.
int extract_labeled_pattern(Khuzdul::ApplyFunction fn, int max_pattern_size, LabelId num_labels, bool root_known,
                            LabelId root_label, LabeledPlan* out, std::string* error);

// Called by the out-of-line VertexSet operations: intersect_with /
// subtracted_by, their bounded forms (result ids below upper_bound) and
// intersect_and_count (returns the count the probe's content gives).
void probe_set_op(int op, const VertexSet* lhs, const VertexSet* rhs, VertexSet* result);
void probe_set_op(int op, const VertexSet* lhs, const VertexSet* rhs, VertexSet* result, VertexId upper_bound);
VertexId probe_count_op(const VertexSet* lhs, const VertexSet* rhs);

}  // namespace kzk

#endif
