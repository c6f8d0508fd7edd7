// Device helpers shared by the kernels of libkz_hip.so.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
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
__device__ __forceinline__ uint32_t hash_slot(uint32_t x, uint32_t shift) { return (x * 2654435761u) >> shift; }

__device__ __forceinline__ void hash_insert(uint32_t* keys, uint16_t* vals, uint32_t mask, uint32_t shift,
                                            uint32_t x, uint16_t idx) {
    uint32_t h = hash_slot(x, shift);
    while (true) {
        const uint32_t old = atomicCAS(&keys[h], kEmpty, x);
        if (old == kEmpty) {
            vals[h] = idx;
            return;
        }
        h = (h + 1) & mask;
    }
}

// Continues a probe for x at slot h.
__device__ __forceinline__ uint32_t hash_probe(const uint32_t* keys, const uint16_t* vals, uint32_t mask, uint32_t h,
                                               uint32_t x) {
    while (true) {
        const uint32_t k = keys[h];
        if (k == x) return vals[h];
        if (k == kEmpty) return kNotFound;
        h = (h + 1) & mask;
    }
}

__device__ __forceinline__ uint32_t hash_find(const uint32_t* keys, const uint16_t* vals, uint32_t mask,
                                              uint32_t shift, uint32_t x) {
    return hash_probe(keys, vals, mask, hash_slot(x, shift), x);
}

// Persistent kernels pull tasks from a device cursor in chunks: one atomic
// reserves `size` consecutive tasks.  Tasks come largest first, so the size
// ramps up with the tasks already taken (at most base / (ramp * workers))
// and shrinks again with the tasks left (at most an eighth of the caller's
// share of them), within [floor, cap]: the heavy first tasks still go one at
// a time, the many light ones in chunks.  A slow ramp (8), floor 1 and cap
// kMaxTaskChunk suit kernels whose task costs vary by orders of magnitude
// (workgroup tasks); kernels of small, similar tasks (one root per wave)
// take fixed chunks (floor = cap), since their last tasks would otherwise
// queue on the cursor one atomic each.  Each cursor
// is the first u32 of its own kCursorStride-word (256-byte) line: atomics on
// one line are serialised at one L2 channel, on the order of 10^8 per
// second, which bounds a kernel of short tasks taking one per atomic.
constexpr uint32_t kCursorStride = 64;
constexpr uint32_t kMaxTaskChunk = 32;

__device__ __forceinline__ uint32_t task_chunk(uint32_t base, uint32_t ntasks, uint32_t workers, uint32_t ramp,
                                               uint32_t floor, uint32_t cap) {
    const uint32_t left = base < ntasks ? ntasks - base : 0u;
    const uint32_t a = left / (8u * workers), b = base / (ramp * workers);
    const uint32_t c = a < b ? a : b;
    return c < floor ? floor : (c > cap ? cap : c);
}

struct TaskChunk {
    uint32_t next, end, size, ntasks, workers, ramp, floor, cap;
    __device__ TaskChunk(uint32_t n, uint32_t w, uint32_t r, uint32_t f = 1, uint32_t c = kMaxTaskChunk)
        : next(0), end(0), size(f), ntasks(n), workers(w), ramp(r), floor(f), cap(c) {}
    __device__ void take(uint32_t base) {
        next = base;
        end = base + size;
        size = task_chunk(end, ntasks, workers, ramp, floor, cap);
    }
};

// Lane 0's value as a wave-uniform value (an SGPR: the compiler then keeps
// what follows from it scalar; __shfl(v, 0) reads as per-lane and turns the
// loops it bounds into divergent ones).  Call with every lane active.
__device__ __forceinline__ uint32_t wave_bcast(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }
// Lane l's value (l wave-uniform) as a wave-uniform value.
__device__ __forceinline__ uint32_t wave_lane(uint32_t v, uint32_t l) { return __builtin_amdgcn_readlane(v, l); }
__device__ __forceinline__ uint64_t wave_lane(uint64_t v, uint32_t l) {
    return (uint64_t(wave_lane(uint32_t(v >> 32), l)) << 32) | wave_lane(uint32_t(v), l);
}

// Next task of a wave (call with every lane; the result is wave-uniform,
// >= ntasks when none is left).
__device__ __forceinline__ uint32_t wave_next_task(TaskChunk& c, uint32_t* cursor) {
    if (c.next == c.end) {
        uint32_t v = 0;
        if ((threadIdx.x & 63) == 0) v = atomicAdd(cursor, c.size);
        c.take(wave_bcast(v));
    }
    return c.next++;
}

// Next task of a workgroup, called by one thread only.
__device__ __forceinline__ uint32_t thread_next_task(TaskChunk& c, uint32_t* cursor) {
    if (c.next == c.end) c.take(atomicAdd(cursor, c.size));
    return c.next++;
}

// Load through a pointer known to point at device (HBM) memory.  Graph
// pointers read from query tables are generic to the compiler, which would
// otherwise emit FLAT instructions; those wait on the LDS counter as well
// and serialise latency hiding.  The cast lets it emit global loads.
#define KZ_GLOBAL_AS __attribute__((address_space(1)))
template <typename T>
__device__ __forceinline__ T ldg(const T* p, uint64_t i) {
    return ((const KZ_GLOBAL_AS T*)p)[i];
}

// Binary search for x in a sorted global list.
__device__ __forceinline__ bool bsearch_global(const uint32_t* __restrict__ a, uint32_t n, uint32_t x) {
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        const uint32_t v = ldg(a, mid);
        if (v < x) lo = mid + 1;
        else hi = mid;
    }
    return lo < n && ldg(a, lo) == x;
}

// First index in a sorted list with a[i] >= x (lower bound).
__device__ __forceinline__ uint32_t lower_bound_global(const uint32_t* __restrict__ a, uint32_t n, uint32_t x) {
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (ldg(a, mid) < x) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// First index with a[i] > x (upper bound).
__device__ __forceinline__ uint32_t upper_bound_global(const uint32_t* __restrict__ a, uint32_t n, uint32_t x) {
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (ldg(a, mid) <= x) lo = mid + 1;
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
// a degree-oriented graph.
struct CliqueView {
    const uint64_t* rowptr = nullptr;
    const uint32_t* col = nullptr;
    std::vector<uint32_t> h_fdeg;
    DeviceGraph* owner = nullptr;
};

// Degree-bucket bounds on |N+(u)| (last bucket unbounded), block sizes and
// the launch policy (serial: bucket kernels one after another).
struct BucketCfg {
    uint32_t hi[kNumBuckets];
    uint32_t block[kNumBuckets];
    uint32_t serial;
};
BucketCfg default_buckets();
int check_buckets(const BucketCfg& c);
std::vector<BucketCfg> tuning_candidates();

// One counting query of the persistent form: k-cliques of one forward graph
// rooted at the vertices in [root_lo, root_hi).
struct CliqueQuery {
    const uint64_t* rowptr;
    const uint32_t* col;
    const std::vector<uint32_t>* fdeg;  // host forward out-degrees
    int k;
    uint32_t root_lo = 0, root_hi = 0xFFFFFFFFu;
};

// Prepared persistent execution of a set of queries: task groups per
// (k, bucket) and the captured hipGraph that runs them all.
struct CliqueJob {
    struct Group {
        int k = 0;
        int bucket = 0;
        uint64_t first = 0, count = 0;
        uint32_t dmax = 0;
        uint32_t block = 0;  // workgroup size; 0: the bucket's configured one
        int cursor = 0;
    };
    bool ready = false;
    BucketCfg cfg{};
    uint32_t nqueries = 0;
    uint64_t ntasks = 0;
    std::vector<Group> groups;
    void* d_queries = nullptr;  // QueryRef[nqueries]
    uint32_t* d_task_v = nullptr;
    uint32_t* d_task_q = nullptr;  // null for a single query
    uint32_t* d_cursor = nullptr;
    unsigned long long* d_result = nullptr;
    // hub roots (last bucket), per k: split into row chunks over many
    // workgroups; rows go to an HBM slab, then one wave counts each row
    struct HubPlan {
        uint32_t nhubs = 0, dmax = 0, wmax = 0, rpc = 64, nitems = 0, nrows = 0, ntiles = 0;
        bool queues = false;  // rows kernel with probe queues
        uint32_t fkey = 16;   // filter bits per id of a hub's list
    };
    std::map<int, HubPlan> hubplan;
    std::map<int, void*> hub_meta;         // HubMeta[nhubs]
    std::map<int, uint32_t*> hub_items;    // exclusive prefix of row chunks per hub
    std::map<int, uint32_t*> hub_rows;     // exclusive prefix of rows per hub
    std::map<int, uint32_t*> hub_tiles;    // exclusive prefix of 32 x 32 count tiles per hub (k = 4)
    std::map<int, unsigned long long*> slabs;  // per k, HBM bitmap rows
    uint64_t bytes = 0;                        // device memory this job holds
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    ~CliqueJob();
};

// Forward part (entries > v) of sorted symmetric lists: per vertex the index
// of the first entry > v and the forward degree; then a wave per vertex
// copies the suffixes (kz_count.hip).
__global__ void fwd_start_kernel(const uint64_t* __restrict__ rowptr, const uint32_t* __restrict__ col, uint32_t n,
                                 uint32_t* __restrict__ skip, uint32_t* __restrict__ fdeg);
__global__ void fwd_copy_kernel(const uint64_t* __restrict__ rowptr, const uint32_t* __restrict__ col, uint32_t n,
                                const uint32_t* __restrict__ skip, const uint64_t* __restrict__ frowptr,
                                uint32_t* __restrict__ fcol);

int clique_job_prepare(CliqueJob& x, const std::vector<CliqueQuery>& queries, const BucketCfg& cfg);
int clique_job_run(CliqueJob& x, uint64_t* counts, double* device_ms);
// Launches a prepared job on stream s and copies its per-query counts to
// host_results (pinned, nqueries entries) on the same stream; no wait.
int clique_job_launch(CliqueJob& x, hipStream_t s, unsigned long long* host_results);
// Prepared single-graph job for (graph, k), built on first use and kept in
// the graph (kz_count.hip).
int graph_job(DeviceGraph& d, int k, CliqueJob** out);
// Motif census of a symmetric resident graph; the relabelled copy and
// counter slabs are kept in the graph (kz_census.hip).  census_bytes is the
// device memory that state holds (0 before the first call).
int motif_census(DeviceGraph& d, uint64_t* counts, kz_run_stats* stats);
uint64_t census_bytes(const DeviceGraph& d);
int clique_workload(const CliqueView& gv, uint64_t* out6);
// Forward part of a graph for the clique kernels: the graph itself when it
// is degree-oriented, else its orientation by (degree, id) rank, built once
// and kept in the graph (fwd_bytes of device memory).
int ensure_forward(DeviceGraph& g);
// Frees the forward copy of a symmetric graph (no-op for oriented graphs);
// no clique execution on the graph may be left.
void release_forward(DeviceGraph& g);
// Degree orientation of a resident symmetric graph (kz_gpugen.hip): ids
// become ranks by (degree, id), u -> v kept for rank v > rank u, lists
// sorted; outputs written only on success (deg_out may be null).
int orient_resident(const DeviceGraph& g, uint64_t** rowptr_out, uint32_t** col_out, uint32_t** deg_out,
                    std::vector<uint32_t>* h_deg, uint64_t* nnz_out);

}  // namespace kz
