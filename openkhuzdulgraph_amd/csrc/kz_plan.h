// Internal view of the plan IR (kz_plan_desc) used by the executors.
#pragma once

#include <cstdint>

#include "kz_device.h"

namespace kz {

typedef kz_plan_desc PlanDesc;

// Fills a kautomine k-clique plan (k = 3..5).
int plan_clique(int k, PlanDesc* out);
int plan_validate(const PlanDesc& p);

// Plain level-by-level execution of a plan on a resident graph.
int plan_execute(DeviceGraph& g, const PlanDesc& p, bool count_only_last, uint64_t* raw_counts, kz_run_stats* st);

// Order code of matched vertices m[0..s) (dense ranks, base s digits).
static inline __host__ __device__ uint32_t order_code(const uint32_t* m, int s) {
    uint32_t code = 0, pw = 1;
    for (int j = 0; j < s; ++j) {
        uint32_t r = 0;
        // dense rank = number of distinct values < m[j]
        for (int i = 0; i < s; ++i) {
            if (m[i] < m[j]) {
                bool first = true;
                for (int k = 0; k < i; ++k)
                    if (m[k] == m[i]) first = false;
                r += first;
            }
        }
        code += r * pw;
        pw *= uint32_t(s);
    }
    return code;
}

extern uint64_t g_level_budget;

}  // namespace kz
