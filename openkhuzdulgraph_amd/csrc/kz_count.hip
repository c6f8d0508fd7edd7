// C-ABI entry points for pattern counting on a resident graph.
//
// kz_count_cliques replaces DistributedApplication::run
// (src/distributed_application.cc:259-356) + Aggregator<EdgeId>::evaluate
// (include/engine.h:628-645) for the clique apps: it returns the same uint64
// that the app prints after the multiplicity division (1 for cliques).

#include <algorithm>
#include <chrono>
#include <cstring>

#include "kz_device.h"
#include "kz_kernels.h"
#include "kz_plan.h"

namespace kz {

// For each vertex, the index of the first neighbour > v (symmetric graphs).
__global__ void fwd_start_kernel(const uint64_t* __restrict__ rowptr, const uint32_t* __restrict__ col, uint32_t n,
                                 uint32_t* __restrict__ skip, uint32_t* __restrict__ fdeg) {
    const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n) return;
    const uint64_t b = rowptr[v];
    const uint32_t d = uint32_t(rowptr[v + 1] - b);
    const uint32_t s = upper_bound_global(col + b, d, v);
    skip[v] = s;
    fdeg[v] = d - s;
}

// Wave per vertex copies the forward part of each list.
__global__ void fwd_copy_kernel(const uint64_t* __restrict__ rowptr, const uint32_t* __restrict__ col, uint32_t n,
                                const uint32_t* __restrict__ skip, const uint64_t* __restrict__ frowptr,
                                uint32_t* __restrict__ fcol) {
    const uint32_t v = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const uint32_t lane = threadIdx.x & 63;
    if (v >= n) return;
    const uint64_t src = rowptr[v] + skip[v];
    const uint64_t dst = frowptr[v];
    const uint32_t len = uint32_t(frowptr[v + 1] - dst);
    for (uint32_t i = lane; i < len; i += 64) fcol[dst + i] = col[src + i];
}

int ensure_forward(DeviceGraph& g) {
    if (g.fwd_ready) return KZ_OK;
    if (g.oriented) {
        g.fwd_rowptr = g.rowptr;
        g.fwd_col = g.col;
        g.h_fdeg = g.h_deg;
        g.fwd_ready = true;
        return KZ_OK;
    }
    hipStream_t s = rt().stream;
    uint32_t *skip = nullptr, *fdeg = nullptr;
    KZ_HIP(hipMalloc(&skip, sizeof(uint32_t) * std::max<uint32_t>(1, g.n)));
    KZ_HIP(hipMalloc(&fdeg, sizeof(uint32_t) * std::max<uint32_t>(1, g.n)));
    if (g.n) hipLaunchKernelGGL(fwd_start_kernel, dim3((g.n + 255) / 256), dim3(256), 0, s, g.rowptr, g.col, g.n, skip, fdeg);
    g.h_fdeg.resize(g.n);
    KZ_HIP(hipMemcpyAsync(g.h_fdeg.data(), fdeg, sizeof(uint32_t) * g.n, hipMemcpyDeviceToHost, s));
    KZ_HIP(hipStreamSynchronize(s));
    std::vector<uint64_t> frp(size_t(g.n) + 1, 0);
    for (uint32_t v = 0; v < g.n; ++v) frp[v + 1] = frp[v] + g.h_fdeg[v];
    KZ_HIP(hipMalloc(&g.fwd_rowptr, sizeof(uint64_t) * frp.size()));
    KZ_HIP(hipMalloc(&g.fwd_col, sizeof(uint32_t) * std::max<uint64_t>(1, frp[g.n])));
    KZ_HIP(hipMemcpyAsync(g.fwd_rowptr, frp.data(), sizeof(uint64_t) * frp.size(), hipMemcpyHostToDevice, s));
    if (g.n)
        hipLaunchKernelGGL(fwd_copy_kernel, dim3((uint64_t(g.n) * 64 + 255) / 256), dim3(256), 0, s, g.rowptr, g.col,
                           g.n, skip, g.fwd_rowptr, g.fwd_col);
    KZ_HIP(hipStreamSynchronize(s));
    hipFree(skip);
    hipFree(fdeg);
    g.fwd_ready = true;
    return KZ_OK;
}

}  // namespace kz

using namespace kz;

extern "C" {

int kz_count_cliques(kz_graph g, int k, int method, uint64_t* count, kz_run_stats* stats) {
    if (!g || !count) KZ_FAIL(KZ_ERR_INVALID, "null argument");
    if (k < 3 || k > 5) KZ_FAIL(KZ_ERR_UNSUPPORTED, "clique size %d outside 3..5 (MAX_PATTERN_SIZE 5)", k);
    if (int e = ensure_ready()) return e;
    auto t0 = std::chrono::steady_clock::now();
    DeviceGraph& d = g->g;
    if (int e = ensure_forward(d)) return e;
    double dev_ms = 0;
    kz_run_stats st;
    std::memset(&st, 0, sizeof(st));
    if (method == KZ_METHOD_PERSISTENT) {
        auto& slot = d.clique_exec[k];
        if (!slot) slot = std::shared_ptr<void>(new CliquePlanExec(), [](void* p) { delete static_cast<CliquePlanExec*>(p); });
        CliquePlanExec& x = *static_cast<CliquePlanExec*>(slot.get());
        CliqueView v;
        v.rowptr = d.fwd_rowptr;
        v.col = d.fwd_col;
        v.h_fdeg = d.h_fdeg;
        v.owner = &d;
        if (int e = clique_prepare(v, k, x)) {
            slot.reset();
            return e;
        }
        if (int e = clique_run(x, count, &dev_ms)) return e;
        st.launches = 1;
        st.tasks = x.ntasks;
    } else if (method == KZ_METHOD_PLAIN) {
        PlanDesc pd;
        plan_clique(k, &pd);
        uint64_t counts[KZ_MAX_AGGREGATORS] = {0};
        if (int e = plan_execute(d, pd, true, counts, &st)) return e;
        *count = counts[0];
        dev_ms = st.device_ms;
    } else {
        KZ_FAIL(KZ_ERR_INVALID, "unknown method %d", method);
    }
    auto t1 = std::chrono::steady_clock::now();
    st.device_ms = dev_ms;
    st.wall_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
    if (stats) *stats = st;
    return KZ_OK;
}

int kz_clique_workload(kz_graph g, uint64_t* out6) {
    if (!g || !out6) KZ_FAIL(KZ_ERR_INVALID, "null argument");
    if (int e = ensure_ready()) return e;
    DeviceGraph& d = g->g;
    if (int e = ensure_forward(d)) return e;
    CliqueView v;
    v.rowptr = d.fwd_rowptr;
    v.col = d.fwd_col;
    v.h_fdeg = d.h_fdeg;
    v.owner = &d;
    return clique_workload(v, out6);
}

}  // extern "C"
