// Device-side shared declarations for libkz_hip.so (gfx950 / CDNA4).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <memory>
#include <vector>

#include "kz_internal.h"

// On failure the runtime's last-error state is cleared, so that a failed
// allocation is not reported again by an unrelated later call.
#define KZ_HIP(call)                                                                   \
    do {                                                                               \
        hipError_t _e = (call);                                                        \
        if (_e != hipSuccess) {                                                        \
            (void)hipGetLastError();                                                   \
            KZ_FAIL(KZ_ERR_DEVICE, "%s failed: %s (%s:%d)", #call, hipGetErrorString(_e), \
                    __FILE__, __LINE__);                                               \
        }                                                                              \
    } while (0)

namespace kz {

constexpr int kWave = 64;  // CDNA wavefront width

// Library-wide device state (one device per process).
struct Runtime {
    int device = -1;
    hipStream_t stream = nullptr;
    int num_cus = 0;
    size_t lds_per_block = 0;
    bool ready = false;
};
Runtime& rt();
int ensure_ready();

// Bytes allocated for a device column array of n entries: kernels read
// neighbour lists as aligned 16-byte groups of 4 entries, so the last group
// may extend up to 3 entries past the end (those entries are masked).
inline size_t col_alloc_bytes(uint64_t n) { return sizeof(uint32_t) * (((n + 3) & ~uint64_t(3)) + 4); }

// A CSR resident in HBM.  rowptr is u64 (graphs beyond 2^32 entries),
// col and deg are u32.  `vbucket` holds vertex ids grouped by out-degree
// bucket for the bitmap kernels (built lazily, see kz_clique.hip).
struct DeviceGraph {
    uint32_t n = 0;
    uint64_t nnz = 0;
    uint32_t max_degree = 0;
    int oriented = 0;
    uint64_t* rowptr = nullptr;
    uint32_t* col = nullptr;
    uint32_t* deg = nullptr;
    // host copies of the degree array (for bucketing and plan sizing)
    std::vector<uint32_t> h_deg;
    uint64_t bytes = 0;
    // forward part of a symmetric graph for the clique kernels, built on
    // first use: vertices renumbered by (degree, id) rank, u -> v kept for
    // rank v > rank u (fwd_bytes of device memory); for oriented graphs
    // these alias rowptr/col.
    uint64_t* fwd_rowptr = nullptr;
    uint32_t* fwd_col = nullptr;
    std::vector<uint32_t> h_fdeg;
    bool fwd_ready = false;
    uint64_t fwd_bytes = 0;
    uint32_t fwd_max = 0;  // largest forward out-degree
    // prepared persistent clique executions, indexed by k (3..5), and the
    // degree-bucket configuration each uses (set or autotuned)
    std::shared_ptr<void> clique_exec[6];
    // motif-census state: the graph relabelled by (degree, id) rank, its
    // forward part, counter slabs and the 4-clique job (kz_census.hip)
    std::shared_ptr<void> census;
    // vertex labels and the label-sorted lists of frequent subgraph mining
    // (kz_fsm.hip), set by kz_graph_set_labels
    std::shared_ptr<void> fsm;
    bool bucket_set[6] = {};
    uint32_t bucket_hi[6][5] = {};
    uint32_t bucket_block[6][5] = {};
    uint32_t bucket_serial[6] = {};
};

// Device scratch arena that grows on demand and is reused across calls.
struct Scratch {
    void* ptr = nullptr;
    size_t size = 0;
};
int scratch_reserve(Scratch& s, size_t bytes);
int fsm_set_labels(DeviceGraph& d, const uint32_t* labels);

}  // namespace kz

struct kz_graph_s {
    kz::DeviceGraph g;
};
