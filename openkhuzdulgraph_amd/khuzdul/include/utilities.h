// Search helpers used by the generated apps (include/utilities.h:64-86 of
// the reference): trim_v_list returns the number of leading ids below an
// upper bound, trim_v_list_lower_bound the index of the first id above a
// lower bound.  Same results; during plan probing the list is reported to
// the probe so the count that follows is attributed to it.
#ifndef UTILITIES_H
#define UTILITIES_H

#include <assert.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <string>

#include "kz_probe_hooks.h"
#include "types.h"

inline VertexId trim_v_list(VertexId* v_list, VertexId num_v, VertexId upper_bound) {
    if (__builtin_expect(kzk::g_hooks != nullptr, 0)) kzk::g_hooks->on_trim(v_list, num_v);
    return VertexId(std::lower_bound(v_list, v_list + num_v, upper_bound) - v_list);
}

inline VertexId trim_v_list_lower_bound(VertexId* v_list, VertexId num_v, VertexId lower_bound) {
    if (__builtin_expect(kzk::g_hooks != nullptr, 0)) kzk::g_hooks->on_trim(v_list, num_v);
    return VertexId(std::upper_bound(v_list, v_list + num_v, lower_bound) - v_list);
}

bool file_exits(const std::string& path);
long file_size(const char* path);

#endif
