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

// Called by the out-of-line VertexSet operations.
void probe_set_op(int op, const VertexSet* lhs, const VertexSet* rhs, VertexSet* result);

}  // namespace kzk

#endif
