// Device helpers shared by the kernels of libkz_hip.so.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <memory>
#include <vector>

#include "kz_device.h"

namespace kz {

constexpr uint32_t kEmpty = 0xFFFFFFFFu;     // hash slot marker (ids are < 2^32-1)
constexpr uint32_t kNotFound = 0xFFFFFFFFu;  // hash miss
constexpr int kNumBuckets = 5;               // degree buckets of the bitmap kernels
constexpr uint32_t kMaxBitmapDegree = 8192;  // largest N+(u) handled by bitmaps
constexpr int kStaticLdsReserve = 1024;      // static LDS headroom of the block kernels

// Open-addressing LDS hash, vertex id -> local index.  Multiplicative
// hashing on the top bits; linear probing; the table is at least twice the
// key count so probes stay short.
__device__ __forceinline__ void hash_insert(uint32_t* keys, uint16_t* vals, uint32_t mask, uint32_t shift,
                                            uint32_t x, uint16_t idx) {
    uint32_t h = (x * 2654435761u) >> shift;
    while (true) {
        const uint32_t old = atomicCAS(&keys[h], kEmpty, x);
        if (old == kEmpty) {
            vals[h] = idx;
            return;
        }
        h = (h + 1) & mask;
    }
}

__device__ __forceinline__ uint32_t hash_find(const uint32_t* keys, const uint16_t* vals, uint32_t mask,
                                              uint32_t shift, uint32_t x) {
    uint32_t h = (x * 2654435761u) >> shift;
    while (true) {
        const uint32_t k = keys[h];
        if (k == x) return vals[h];
        if (k == kEmpty) return kNotFound;
        h = (h + 1) & mask;
    }
}

// Binary search for x in a sorted global list.
__device__ __forceinline__ bool bsearch_global(const uint32_t* __restrict__ a, uint32_t n, uint32_t x) {
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        const uint32_t v = a[mid];
        if (v < x) lo = mid + 1;
        else hi = mid;
    }
    return lo < n && a[lo] == x;
}

// First index in a sorted list with a[i] >= x (lower bound).
__device__ __forceinline__ uint32_t lower_bound_global(const uint32_t* __restrict__ a, uint32_t n, uint32_t x) {
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (a[mid] < x) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// First index with a[i] > x (upper bound).
__device__ __forceinline__ uint32_t upper_bound_global(const uint32_t* __restrict__ a, uint32_t n, uint32_t x) {
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (a[mid] <= x) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// Position of the n-th (0-based) set bit of w (n < popcount(w)).
__device__ __forceinline__ uint32_t select_bit(unsigned long long w, uint32_t n) {
    uint32_t pos = 0;
    uint32_t c = __popc(uint32_t(w));
    if (n >= c) {
        n -= c;
        w >>= 32;
        pos += 32;
    }
    uint32_t x = uint32_t(w);
    c = __popc(x & 0xFFFFu);
    if (n >= c) {
        n -= c;
        x >>= 16;
        pos += 16;
    }
    c = __popc(x & 0xFFu);
    if (n >= c) {
        n -= c;
        x >>= 8;
        pos += 8;
    }
    c = __popc(x & 0xFu);
    if (n >= c) {
        n -= c;
        x >>= 4;
        pos += 4;
    }
    c = __popc(x & 0x3u);
    if (n >= c) {
        n -= c;
        x >>= 2;
        pos += 2;
    }
    c = x & 1u;
    if (n >= c) pos += 1;
    return pos;
}

// Make this wave's LDS writes visible to its other lanes before they read
// (waits on the LDS counter; waves of a workgroup may be elsewhere in their
// loops, so no workgroup barrier is used).
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

// Sum of a uint64 across the 64 lanes of a wave (butterfly over DPP/ds
// swizzles emitted for __shfl_xor).
__device__ __forceinline__ unsigned long long wave_sum_u64(unsigned long long v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

// Graph view the clique kernels run on: the forward (u -> v, v > u) part of
// the graph, which is the graph itself when it is degree-oriented.
struct CliqueView {
    const uint64_t* rowptr = nullptr;
    const uint32_t* col = nullptr;
    std::vector<uint32_t> h_fdeg;
    DeviceGraph* owner = nullptr;
};

// Prepared persistent execution for one (graph, k).
struct CliquePlanExec {
    bool ready = false;
    std::vector<uint64_t> off;  // bucket offsets into d_tasks
    uint64_t ntasks = 0;
    uint32_t max_task_deg = 0;
    uint32_t slab_blocks = 0;
    uint32_t* d_tasks = nullptr;
    uint32_t* d_cursor = nullptr;
    unsigned long long* d_result = nullptr;
    unsigned long long* d_slab = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    ~CliquePlanExec();
};

int clique_prepare(const CliqueView& gv, int k, CliquePlanExec& x);
int clique_run(CliquePlanExec& x, uint64_t* count, double* device_ms);
int clique_workload(const CliqueView& gv, uint64_t* out6);
int ensure_forward(DeviceGraph& g);

}  // namespace kz
