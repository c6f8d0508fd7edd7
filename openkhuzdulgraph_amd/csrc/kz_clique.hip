// k-clique counting (k = 3, 4, 5): the batched / persistent form.
//
// Reference path replaced: the kautomine/kgraphpi clique schedules
// (kautomine/{triangle,four_clique,five_clique}.cc apply(), kgraphpi/ same
// names) executed by EmbeddingExplorationEngine::extend_embeddings
// (src/engine.cc:914-1118) with VertexSet::intersect_with
// (src/vertex_set.cc:20-49) at every level.  Every one of those schedules
// counts each k-clique once, at its lowest-id vertex v_0 with
// v_0 < v_1 < ... < v_{k-1}: on a degree-oriented graph (ids = degree
// ranks, every stored edge u->v has v > u) that is the number of
// (k-1)-cliques inside the out-neighbourhood N+(u).
//
// MI355X design (no counterpart in the reference):
//   * one task = one root vertex u.  N+(u) is staged in LDS together with an
//     open-addressing hash (vertex id -> local index);
//   * the induced DAG on N+(u) is built as bitmap rows in LDS: row i holds
//     bit j (j > i) iff N+(u)[j] is in N+(N+(u)[i]).  Each row is one
//     coalesced stream over N+(N+(u)[i]) with hash probes; for skewed pairs
//     (|N+(v)| much longer than the remaining candidates) the candidates are
//     binary-searched in N+(v) instead;
//   * cliques are then counted with wave64 AND + popcount over the rows:
//     k=4 sums popc(row_i & row_j) over DAG edges (i, j); k=5 descends one
//     level more; k=3 counts the row bits directly;
//   * tasks are bucketed by |N+(u)|: <= 64 runs one task per wave (one u64
//     row per candidate), up to 1024 runs one task per workgroup with the
//     rows in LDS, larger ones keep rows in a per-workgroup HBM slab;
//   * each bucket is a persistent kernel pulling tasks (largest first) from
//     a device atomic cursor; all buckets are captured in one hipGraph with
//     parallel branches and replayed per query.
// Per-wave uint64 partial counts are reduced in LDS and added with one
// 64-bit atomic per workgroup.

#include <algorithm>
#include <chrono>
#include <cstring>

#include "kz_device.h"
#include "kz_kernels.h"

namespace kz {

// ------------------------------------------------------------ wave tasks --
// One task per wave, |N+(u)| <= 64.  LDS per wave: L[64] ids, 128-slot hash,
// 64 bitmap rows (one u64 each).
template <int K>
__global__ __launch_bounds__(256) void clique_wave_kernel(const uint64_t* __restrict__ rowptr,
                                                          const uint32_t* __restrict__ col,
                                                          const uint32_t* __restrict__ tasks, uint32_t ntasks,
                                                          uint32_t* __restrict__ cursor,
                                                          unsigned long long* __restrict__ result) {
    constexpr int WPB = 4;
    constexpr uint32_t H = 128;
    __shared__ uint32_t sL[WPB][64];
    __shared__ uint32_t sKey[WPB][H];
    __shared__ uint16_t sVal[WPB][H];
    __shared__ unsigned long long sRow[WPB][64];
    __shared__ unsigned long long sRed[WPB];
    const int lane = threadIdx.x & 63;
    const int w = threadIdx.x >> 6;
    uint32_t* L = sL[w];
    uint32_t* hk = sKey[w];
    uint16_t* hv = sVal[w];
    unsigned long long* rows = sRow[w];
    unsigned long long acc = 0;
    while (true) {
        uint32_t t = 0;
        if (lane == 0) t = atomicAdd(cursor, 1u);
        t = __shfl(t, 0);
        if (t >= ntasks) break;
        const uint32_t u = tasks[t];
        const uint64_t ub = rowptr[u];
        const uint32_t d = uint32_t(rowptr[u + 1] - ub);
        // stage N+(u), clear hash and rows
        uint32_t x = lane < d ? col[ub + lane] : 0xFFFFFFFFu;
        L[lane] = x;
        hk[lane] = kEmpty;
        hk[lane + 64] = kEmpty;
        rows[lane] = 0;
        wave_lds_sync();
        if (lane < d) hash_insert(hk, hv, H - 1, 25, x, uint16_t(lane));
        wave_lds_sync();
        const uint32_t lmax = L[d - 1];
        // build rows (or count triangles directly for K == 3)
        for (uint32_t i = 0; i + 1 < d; ++i) {
            const uint32_t v = L[i];
            const uint64_t vb = rowptr[v];
            const uint32_t dv = uint32_t(rowptr[v + 1] - vb);
            const uint32_t rem = d - 1 - i;
            if (dv > 16u * rem) {
                const uint32_t j = i + 1 + lane;
                if (j < d && bsearch_global(col + vb, dv, L[j])) {
                    if (K == 3) acc += 1;
                    else atomicOr(&rows[i], 1ull << j);
                }
            } else {
                for (uint32_t base = 0; base < dv; base += 64) {
                    const uint32_t p = base + lane;
                    const uint32_t y = p < dv ? col[vb + p] : 0xFFFFFFFFu;
                    if (y > v && y <= lmax) {
                        const uint32_t j = hash_find(hk, hv, H - 1, 25, y);
                        if (j != kNotFound) {
                            if (K == 3) acc += 1;
                            else atomicOr(&rows[i], 1ull << j);
                        }
                    }
                    if (__ballot(y > lmax)) break;
                }
            }
        }
        wave_lds_sync();
        if (K >= 4 && lane < d) {
            const unsigned long long ri = rows[lane];
            unsigned long long m = ri;
            while (m) {
                const int j = __builtin_ctzll(m);
                m &= m - 1;
                const unsigned long long m2 = ri & rows[j];
                if (K == 4) {
                    acc += __popcll(m2);
                } else {
                    unsigned long long m3 = m2;
                    while (m3) {
                        const int l = __builtin_ctzll(m3);
                        m3 &= m3 - 1;
                        acc += __popcll(m2 & rows[l]);
                    }
                }
            }
        }
        wave_lds_sync();
    }
    acc = wave_sum_u64(acc);
    if (lane == 0) sRed[w] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long s = 0;
        for (int i = 0; i < WPB; ++i) s += sRed[i];
        if (s) atomicAdd(result, s);
    }
}

// ----------------------------------------------------------- block tasks --
// One task per workgroup.  Dynamic LDS: [rows Dmax*W u64 (when rows are in
// LDS)][L Dmax u32][hash keys H u32][hash vals H u16].  With GLOBAL_ROWS the
// rows live in an HBM slab of Dmax*W u64 per workgroup.
template <int K, bool GLOBAL_ROWS>
__global__ __launch_bounds__(512) void clique_block_kernel(const uint64_t* __restrict__ rowptr,
                                                           const uint32_t* __restrict__ col,
                                                           const uint32_t* __restrict__ tasks, uint32_t ntasks,
                                                           uint32_t* __restrict__ cursor,
                                                           unsigned long long* __restrict__ result, uint32_t Dmax,
                                                           uint32_t W, uint32_t H, uint32_t hshift,
                                                           unsigned long long* __restrict__ gslab) {
    extern __shared__ __align__(16) unsigned char smem[];
    unsigned long long* rows = nullptr;
    unsigned char* p = smem;
    if (GLOBAL_ROWS) {
        rows = gslab + size_t(blockIdx.x) * Dmax * W;
    } else if (K >= 4) {
        rows = reinterpret_cast<unsigned long long*>(p);
        p += size_t(Dmax) * W * 8;
    }
    uint32_t* L = reinterpret_cast<uint32_t*>(p);
    p += size_t(Dmax) * 4;
    uint32_t* hk = reinterpret_cast<uint32_t*>(p);
    p += size_t(H) * 4;
    uint16_t* hv = reinterpret_cast<uint16_t*>(p);
    __shared__ uint32_t sTask;
    __shared__ unsigned long long sRed[16];
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int nwaves = blockDim.x >> 6;
    const uint32_t hmask = H - 1;
    unsigned long long acc = 0;
    while (true) {
        if (threadIdx.x == 0) sTask = atomicAdd(cursor, 1u);
        __syncthreads();
        const uint32_t t = sTask;
        if (t >= ntasks) break;
        const uint32_t u = tasks[t];
        const uint64_t ub = rowptr[u];
        const uint32_t d = uint32_t(rowptr[u + 1] - ub);
        const uint32_t wd = (d + 63) >> 6;
        for (uint32_t i = threadIdx.x; i < d; i += blockDim.x) L[i] = col[ub + i];
        for (uint32_t i = threadIdx.x; i < H; i += blockDim.x) hk[i] = kEmpty;
        if (K >= 4)
            for (uint32_t i = threadIdx.x; i < d * W; i += blockDim.x) rows[i] = 0;
        if (GLOBAL_ROWS) __threadfence();
        __syncthreads();
        for (uint32_t i = threadIdx.x; i < d; i += blockDim.x) hash_insert(hk, hv, hmask, hshift, L[i], uint16_t(i));
        __syncthreads();
        const uint32_t lmax = L[d - 1];
        // rows: one wave per row i
        for (uint32_t i = wave; i + 1 < d; i += nwaves) {
            const uint32_t v = L[i];
            const uint64_t vb = rowptr[v];
            const uint32_t dv = uint32_t(rowptr[v + 1] - vb);
            const uint32_t rem = d - 1 - i;
            unsigned long long* ri = rows + size_t(i) * W;
            if (dv > 16u * rem) {
                for (uint32_t j = i + 1 + lane; j < d; j += 64) {
                    if (bsearch_global(col + vb, dv, L[j])) {
                        if (K == 3) acc += 1;
                        else atomicOr(&ri[j >> 6], 1ull << (j & 63));
                    }
                }
            } else {
                for (uint32_t base = 0; base < dv; base += 64) {
                    const uint32_t q = base + lane;
                    const uint32_t y = q < dv ? col[vb + q] : 0xFFFFFFFFu;
                    if (y > v && y <= lmax) {
                        const uint32_t j = hash_find(hk, hv, hmask, hshift, y);
                        if (j != kNotFound) {
                            if (K == 3) acc += 1;
                            else atomicOr(&ri[j >> 6], 1ull << (j & 63));
                        }
                    }
                    if (__ballot(y > lmax)) break;
                }
            }
        }
        if (GLOBAL_ROWS) __threadfence();
        __syncthreads();
        if (K >= 4) {
            // count: one wave per row i, lanes over the set bits j of row i
            for (uint32_t i = wave; i + 2 < d; i += nwaves) {
                const unsigned long long* ri = rows + size_t(i) * W;
                const uint32_t k_lo = (i + 1) >> 6;
                uint32_t total = 0;
                for (uint32_t k = k_lo; k < wd; ++k) total += __popcll(ri[k]);
                for (uint32_t base = 0; base < total; base += 64) {
                    const uint32_t r = base + lane;
                    if (r >= total) break;
                    uint32_t k = k_lo, rem = r;
                    unsigned long long wv = ri[k];
                    for (uint32_t c = __popcll(wv); rem >= c; c = __popcll(wv)) {
                        rem -= c;
                        wv = ri[++k];
                    }
                    const uint32_t j = k * 64 + select_bit(wv, rem);
                    const unsigned long long* rj = rows + size_t(j) * W;
                    const uint32_t kj = (j + 1) >> 6;
                    if (K == 4) {
                        for (uint32_t kk = kj; kk < wd; ++kk) acc += __popcll(ri[kk] & rj[kk]);
                    } else {
                        for (uint32_t kk = kj; kk < wd; ++kk) {
                            unsigned long long m = ri[kk] & rj[kk];
                            while (m) {
                                const uint32_t l = kk * 64 + __builtin_ctzll(m);
                                m &= m - 1;
                                const unsigned long long* rl = rows + size_t(l) * W;
                                for (uint32_t k3 = (l + 1) >> 6; k3 < wd; ++k3)
                                    acc += __popcll(ri[k3] & rj[k3] & rl[k3]);
                            }
                        }
                    }
                }
            }
        }
        __syncthreads();
    }
    acc = wave_sum_u64(acc);
    if (lane == 0) sRed[wave] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long s = 0;
        for (int i = 0; i < nwaves; ++i) s += sRed[i];
        if (s) atomicAdd(result, s);
    }
}

__global__ void reset_u32_kernel(uint32_t* p, int n) {
    if (int(threadIdx.x) < n) p[threadIdx.x] = 0;
}
__global__ void reset_u64_kernel(unsigned long long* p, int n) {
    if (int(threadIdx.x) < n) p[threadIdx.x] = 0;
}

// ------------------------------------------------------------- host side --

// Bucket limits on |N+(u)|.  Bucket 0 is the wave kernel; 1..3 keep rows in
// LDS; 4 keeps rows in HBM.
static const uint32_t kBucketHi[kNumBuckets] = {64, 256, 512, 1024, 0xFFFFFFFFu};

static uint32_t bucket_of(uint32_t d) {
    for (int b = 0; b < kNumBuckets; ++b)
        if (d <= kBucketHi[b]) return uint32_t(b);
    return kNumBuckets - 1;
}

// Vertices with |N+(u)| >= min_deg grouped by bucket, each bucket sorted by
// degree descending (largest tasks first keeps the persistent tail short).
int build_task_lists(DeviceGraph& g, const std::vector<uint32_t>& fdeg, uint32_t min_deg,
                     std::vector<uint32_t>& h_tasks, std::vector<uint64_t>& off) {
    off.assign(kNumBuckets + 1, 0);
    std::vector<std::vector<uint32_t>> b(kNumBuckets);
    for (uint32_t v = 0; v < fdeg.size(); ++v)
        if (fdeg[v] >= min_deg) b[bucket_of(fdeg[v])].push_back(v);
    h_tasks.clear();
    for (int i = 0; i < kNumBuckets; ++i) {
        std::stable_sort(b[i].begin(), b[i].end(), [&](uint32_t x, uint32_t y) { return fdeg[x] > fdeg[y]; });
        off[i] = h_tasks.size();
        h_tasks.insert(h_tasks.end(), b[i].begin(), b[i].end());
    }
    off[kNumBuckets] = h_tasks.size();
    (void)g;
    return KZ_OK;
}

CliquePlanExec::~CliquePlanExec() {
    if (exec) hipGraphExecDestroy(exec);
    if (graph) hipGraphDestroy(graph);
    hipFree(d_tasks);
    hipFree(d_cursor);
    hipFree(d_result);
    hipFree(d_slab);
}

static size_t lds_bytes(uint32_t Dmax, uint32_t W, uint32_t H, bool rows_in_lds) {
    return (rows_in_lds ? size_t(Dmax) * W * 8 : 0) + size_t(Dmax) * 4 + size_t(H) * 4 + size_t(H) * 2 + 16;
}

template <int K>
static int launch_buckets(const CliqueView& gv, CliquePlanExec& x, hipStream_t* streams, int nstreams) {
    const Runtime& R = rt();
    int si = 0;
    for (int b = 0; b < kNumBuckets; ++b) {
        const uint32_t nt = uint32_t(x.off[b + 1] - x.off[b]);
        if (!nt) continue;
        hipStream_t s = streams[(si++) % nstreams];
        const uint32_t* tasks = x.d_tasks + x.off[b];
        uint32_t* cur = x.d_cursor + b;
        if (b == 0) {
            uint32_t grid = std::min<uint32_t>((nt + 3) / 4, uint32_t(R.num_cus) * 8);
            hipLaunchKernelGGL(clique_wave_kernel<K>, dim3(grid), dim3(256), 0, s, gv.rowptr, gv.col, tasks, nt, cur,
                               x.d_result + b);
            continue;
        }
        const uint32_t Dmax = b == kNumBuckets - 1 ? x.max_task_deg : kBucketHi[b];
        const uint32_t W = (Dmax + 63) / 64;
        uint32_t H = 64;
        while (H < 2 * Dmax) H <<= 1;
        uint32_t hbits = 0;
        while ((1u << hbits) < H) ++hbits;
        const uint32_t hshift = 32 - hbits;
        // only k >= 4 builds bitmap rows; the last bucket keeps them in HBM
        const bool global_rows = (b == kNumBuckets - 1) && K >= 4;
        const size_t lds = lds_bytes(Dmax, W, H, !global_rows && K >= 4);
        const uint32_t block = (b >= 3) ? 512 : 256;
        if (lds + kStaticLdsReserve > R.lds_per_block)
            KZ_FAIL(KZ_ERR_UNSUPPORTED, "bucket %d needs %zu bytes of LDS (limit %zu)", b, lds, R.lds_per_block);
        uint32_t per_cu = std::max<uint32_t>(1, uint32_t(R.lds_per_block / (lds + kStaticLdsReserve)));
        per_cu = std::min<uint32_t>(per_cu, 2048 / block);
        uint32_t grid = std::min<uint32_t>(nt, uint32_t(R.num_cus) * per_cu);
        if (global_rows) grid = std::min<uint32_t>(grid, x.slab_blocks);
        if (grid == 0) KZ_FAIL(KZ_ERR_STATE, "empty launch for bucket %d", b);
        if (K >= 4 && !global_rows) {
            hipLaunchKernelGGL((clique_block_kernel<K, false>), dim3(grid), dim3(block), lds, s, gv.rowptr, gv.col,
                               tasks, nt, cur, x.d_result + b, Dmax, W, H, hshift, nullptr);
        } else if (K >= 4) {
            hipLaunchKernelGGL((clique_block_kernel<K, true>), dim3(grid), dim3(block), lds, s, gv.rowptr, gv.col,
                               tasks, nt, cur, x.d_result + b, Dmax, W, H, hshift, x.d_slab);
        } else {
            hipLaunchKernelGGL((clique_block_kernel<K, false>), dim3(grid), dim3(block), lds, s, gv.rowptr, gv.col,
                               tasks, nt, cur, x.d_result + b, Dmax, W, H, hshift, nullptr);
        }
    }
    return KZ_OK;
}

static int set_lds_attrs() {
    static bool done = false;
    if (done) return KZ_OK;
    // Dynamic LDS may use the block limit minus the kernel's static LDS; the
    // attribute is advisory on this runtime, so a refusal is not fatal (the
    // launch itself reports a size it cannot honour).
    const int maxb = int(rt().lds_per_block) - kStaticLdsReserve;
    const void* fns[] = {reinterpret_cast<const void*>(&clique_block_kernel<3, false>),
                         reinterpret_cast<const void*>(&clique_block_kernel<4, false>),
                         reinterpret_cast<const void*>(&clique_block_kernel<5, false>),
                         reinterpret_cast<const void*>(&clique_block_kernel<4, true>),
                         reinterpret_cast<const void*>(&clique_block_kernel<5, true>)};
    for (const void* f : fns)
        if (hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, maxb) != hipSuccess)
            (void)hipGetLastError();
    done = true;
    return KZ_OK;
}

// Prepare (once per graph and k) the task lists, cursors and the captured
// hipGraph with one parallel branch per non-empty bucket.
int clique_prepare(const CliqueView& gv, int k, CliquePlanExec& x) {
    if (x.ready) return KZ_OK;
    if (int e = set_lds_attrs()) return e;
    std::vector<uint32_t> h_tasks;
    if (int e = build_task_lists(*gv.owner, gv.h_fdeg, uint32_t(k - 1), h_tasks, x.off)) return e;
    x.max_task_deg = 0;
    for (uint32_t v : h_tasks) x.max_task_deg = std::max(x.max_task_deg, gv.h_fdeg[v]);
    if (x.max_task_deg > kMaxBitmapDegree)
        KZ_FAIL(KZ_ERR_UNSUPPORTED,
                "persistent clique form needs out-degree <= %u (found %u); orient the graph by degree",
                kMaxBitmapDegree, x.max_task_deg);
    x.ntasks = h_tasks.size();
    KZ_HIP(hipMalloc(&x.d_tasks, sizeof(uint32_t) * std::max<size_t>(1, h_tasks.size())));
    KZ_HIP(hipMalloc(&x.d_cursor, sizeof(uint32_t) * kNumBuckets));
    KZ_HIP(hipMalloc(&x.d_result, sizeof(unsigned long long) * kNumBuckets));
    if (!h_tasks.empty())
        KZ_HIP(hipMemcpy(x.d_tasks, h_tasks.data(), sizeof(uint32_t) * h_tasks.size(), hipMemcpyHostToDevice));
    const uint64_t hub = x.off[kNumBuckets] - x.off[kNumBuckets - 1];
    if (hub && k >= 4) {
        const uint32_t W = (x.max_task_deg + 63) / 64;
        x.slab_blocks = uint32_t(std::min<uint64_t>(hub, uint64_t(rt().num_cus) * 2));
        KZ_HIP(hipMalloc(&x.d_slab, size_t(x.slab_blocks) * x.max_task_deg * W * 8));
    }
    // capture: reset cursors/results, fork one branch per bucket, join
    Runtime& R = rt();
    hipStream_t main = R.stream;
    hipStream_t br[kNumBuckets];
    hipEvent_t fork, join[kNumBuckets];
    for (int i = 0; i < kNumBuckets; ++i) {
        KZ_HIP(hipStreamCreateWithFlags(&br[i], hipStreamNonBlocking));
        KZ_HIP(hipEventCreateWithFlags(&join[i], hipEventDisableTiming));
    }
    KZ_HIP(hipEventCreateWithFlags(&fork, hipEventDisableTiming));
    KZ_HIP(hipStreamBeginCapture(main, hipStreamCaptureModeThreadLocal));
    hipLaunchKernelGGL(reset_u32_kernel, dim3(1), dim3(64), 0, main, x.d_cursor, kNumBuckets);
    hipLaunchKernelGGL(reset_u64_kernel, dim3(1), dim3(64), 0, main, x.d_result, kNumBuckets);
    KZ_HIP(hipEventRecord(fork, main));
    for (int i = 0; i < kNumBuckets; ++i) KZ_HIP(hipStreamWaitEvent(br[i], fork, 0));
    int e = KZ_OK;
    if (k == 3) e = launch_buckets<3>(gv, x, br, kNumBuckets);
    else if (k == 4) e = launch_buckets<4>(gv, x, br, kNumBuckets);
    else e = launch_buckets<5>(gv, x, br, kNumBuckets);
    for (int i = 0; i < kNumBuckets; ++i) {
        KZ_HIP(hipEventRecord(join[i], br[i]));
        KZ_HIP(hipStreamWaitEvent(main, join[i], 0));
    }
    hipGraph_t graph;
    KZ_HIP(hipStreamEndCapture(main, &graph));
    if (e == KZ_OK && hipGetLastError() != hipSuccess) {
        e = KZ_ERR_DEVICE;
        kz::set_error("a bucket launch was rejected during capture");
    }
    if (e) {
        hipGraphDestroy(graph);
        return e;
    }
    x.graph = graph;
    KZ_HIP(hipGraphInstantiate(&x.exec, graph, nullptr, nullptr, 0));
    for (int i = 0; i < kNumBuckets; ++i) {
        KZ_HIP(hipStreamDestroy(br[i]));
        KZ_HIP(hipEventDestroy(join[i]));
    }
    KZ_HIP(hipEventDestroy(fork));
    x.ready = true;
    return KZ_OK;
}

int clique_run(CliquePlanExec& x, uint64_t* count, double* device_ms) {
    Runtime& R = rt();
    hipEvent_t a, b;
    KZ_HIP(hipEventCreate(&a));
    KZ_HIP(hipEventCreate(&b));
    KZ_HIP(hipEventRecord(a, R.stream));
    KZ_HIP(hipGraphLaunch(x.exec, R.stream));
    KZ_HIP(hipEventRecord(b, R.stream));
    unsigned long long res[kNumBuckets];
    KZ_HIP(hipMemcpyAsync(res, x.d_result, sizeof(res), hipMemcpyDeviceToHost, R.stream));
    KZ_HIP(hipStreamSynchronize(R.stream));
    float ms = 0;
    KZ_HIP(hipEventElapsedTime(&ms, a, b));
    hipEventDestroy(a);
    hipEventDestroy(b);
    unsigned long long s = 0;
    for (int i = 0; i < kNumBuckets; ++i) s += res[i];
    *count = s;
    if (device_ms) *device_ms = ms;
    return KZ_OK;
}

// ------------------------------------------------------- workload terms --
// Exact terms of the merge-model byte count of the clique schedules (the
// kautomine four_clique.cc schedule reads N+(u), N+(v) per edge and the
// cached prefix c(u,v) plus N+(w) per triangle).  One workgroup per root u:
// N+(u) in an LDS hash, one wave per v in N+(u) streams N+(v).
//   acc[0] = sum_{(u,v)} c(u,v)            (triangles T)
//   acc[1] = sum_{(u,v)} c(u,v)^2
//   acc[2] = sum_{(u,v,w) in Tri} d+(w)
//   acc[3] = sum_{(u,v)} d+(v)             (= sum_v d-(v) d+(v))
__global__ __launch_bounds__(256) void workload_kernel(const uint64_t* __restrict__ rowptr,
                                                       const uint32_t* __restrict__ col, uint32_t n,
                                                       uint32_t* __restrict__ cursor, uint32_t H, uint32_t hshift,
                                                       unsigned long long* __restrict__ out) {
    extern __shared__ __align__(16) unsigned char smem[];
    uint32_t* hk = reinterpret_cast<uint32_t*>(smem);
    uint16_t* hv = reinterpret_cast<uint16_t*>(smem + size_t(H) * 4);
    __shared__ uint32_t sTask;
    __shared__ unsigned long long sRed[4][4];
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    unsigned long long a0 = 0, a1 = 0, a2 = 0, a3 = 0;
    while (true) {
        if (threadIdx.x == 0) sTask = atomicAdd(cursor, 1u);
        __syncthreads();
        const uint32_t u = sTask;
        if (u >= n) break;
        const uint64_t ub = rowptr[u];
        const uint32_t d = uint32_t(rowptr[u + 1] - ub);
        if (d > 0) {
            for (uint32_t i = threadIdx.x; i < H; i += blockDim.x) hk[i] = kEmpty;
            __syncthreads();
            for (uint32_t i = threadIdx.x; i < d; i += blockDim.x)
                hash_insert(hk, hv, H - 1, hshift, col[ub + i], uint16_t(i & 0xFFFF));
            __syncthreads();
            for (uint32_t i = wave; i < d; i += blockDim.x / 64) {
                const uint32_t v = col[ub + i];
                const uint64_t vb = rowptr[v];
                const uint32_t dv = uint32_t(rowptr[v + 1] - vb);
                uint32_t c = 0;
                unsigned long long sw = 0;
                for (uint32_t base = 0; base < dv; base += 64) {
                    const uint32_t q = base + lane;
                    if (q < dv) {
                        const uint32_t w = col[vb + q];
                        if (hash_find(hk, hv, H - 1, hshift, w) != kNotFound) {
                            c += 1;
                            sw += uint32_t(rowptr[w + 1] - rowptr[w]);
                        }
                    }
                }
                unsigned long long cc = wave_sum_u64(c);
                sw = wave_sum_u64(sw);
                if (lane == 0) {
                    a0 += cc;
                    a1 += cc * cc;
                    a2 += sw;
                    a3 += dv;
                }
            }
        }
        __syncthreads();
    }
    if (lane == 0) {
        sRed[wave][0] = a0;
        sRed[wave][1] = a1;
        sRed[wave][2] = a2;
        sRed[wave][3] = a3;
    }
    __syncthreads();
    if (threadIdx.x < 4) {
        unsigned long long s = 0;
        for (int w = 0; w < 4; ++w) s += sRed[w][threadIdx.x];
        atomicAdd(&out[threadIdx.x], s);
    }
}

int clique_workload(const CliqueView& gv, uint64_t* out6) {
    Runtime& R = rt();
    const uint32_t n = uint32_t(gv.h_fdeg.size());
    uint32_t maxd = 1;
    uint64_t m = 0, sq = 0;
    for (uint32_t v = 0; v < n; ++v) {
        maxd = std::max(maxd, gv.h_fdeg[v]);
        m += gv.h_fdeg[v];
        sq += uint64_t(gv.h_fdeg[v]) * gv.h_fdeg[v];
    }
    if (maxd > kMaxBitmapDegree)
        KZ_FAIL(KZ_ERR_UNSUPPORTED, "workload pass needs forward out-degree <= %u", kMaxBitmapDegree);
    uint32_t H = 64;
    while (H < 2 * maxd) H <<= 1;
    uint32_t bits = 0;
    while ((1u << bits) < H) ++bits;
    const size_t lds = size_t(H) * 6;
    uint32_t* cur = nullptr;
    unsigned long long* acc = nullptr;
    KZ_HIP(hipMalloc(&cur, 4));
    KZ_HIP(hipMalloc(&acc, 32));
    KZ_HIP(hipMemsetAsync(cur, 0, 4, R.stream));
    KZ_HIP(hipMemsetAsync(acc, 0, 32, R.stream));
    if (hipFuncSetAttribute(reinterpret_cast<const void*>(&workload_kernel), hipFuncAttributeMaxDynamicSharedMemorySize,
                            int(R.lds_per_block) - kStaticLdsReserve) != hipSuccess)
        (void)hipGetLastError();
    const uint32_t grid = uint32_t(R.num_cus) * std::max<uint32_t>(1, std::min<uint32_t>(8, uint32_t(R.lds_per_block / (lds + kStaticLdsReserve))));
    hipLaunchKernelGGL(workload_kernel, dim3(grid), dim3(256), lds, R.stream, gv.rowptr, gv.col, n, cur, H, 32 - bits,
                       acc);
    KZ_HIP(hipGetLastError());
    unsigned long long h[4];
    KZ_HIP(hipMemcpyAsync(h, acc, 32, hipMemcpyDeviceToHost, R.stream));
    KZ_HIP(hipStreamSynchronize(R.stream));
    hipFree(cur);
    hipFree(acc);
    out6[0] = m;
    out6[1] = h[0];
    out6[2] = h[1];
    out6[3] = h[2];
    out6[4] = sq;
    out6[5] = h[3];
    return KZ_OK;
}

}  // namespace kz
