// k-clique counting (k = 3, 4, 5): the batched / persistent form.
//
// Reference path replaced: the kautomine/kgraphpi clique schedules
// (kautomine/{triangle,four_clique,five_clique}.cc apply(), kgraphpi/ same
// names) executed by EmbeddingExplorationEngine::extend_embeddings
// (src/engine.cc:914-1118) with VertexSet::intersect_with
// (src/vertex_set.cc:20-49) at every level.  Every one of those schedules
// counts each k-clique once, at its lowest-id vertex v_0 with
// v_0 < v_1 < ... < v_{k-1}: on the forward graph (only u -> v with v > u,
// which is the graph itself after degree orientation) that is the number of
// (k-1)-cliques inside the out-neighbourhood N+(u).
//
// MI355X design (no counterpart in the reference):
//   * one task = one root vertex u of one query (graph, k).  N+(u) is staged
//     in LDS together with an open-addressing hash (vertex id -> index);
//   * the induced DAG on N+(u) is built as bitmap rows: row i holds bit j
//     (j > i) iff N+(u)[j] is in N+(N+(u)[i]).  Each row is one coalesced
//     stream over N+(N+(u)[i]) with hash probes, stopped at max N+(u); for
//     skewed pairs (|N+(v)| much longer than the remaining candidates) the
//     candidates are binary-searched in N+(v) instead;
//   * cliques are counted with wave64 AND + popcount over the rows: k=4 sums
//     popc(row_i & row_j) over DAG edges (i, j); k=5 descends one level more;
//     k=3 counts row bits while they are found;
//   * tasks are bucketed by |N+(u)| with runtime bounds (autotunable): the
//     first bucket (<= 64) runs one task per wave with one u64 row per
//     candidate, the middle buckets one task per workgroup with rows in LDS,
//     the last keeps rows in a per-workgroup HBM slab;
//   * every (k, bucket) is a persistent kernel pulling tasks, largest first,
//     from a device atomic cursor; all of them are captured in one hipGraph
//     with parallel branches, replayed per run.  A batch of many graphs and
//     patterns is the same graph with tasks tagged by query.
// Per-lane uint64 partial counts are reduced per wave and added to the
// query's result with one 64-bit atomic when the wave moves to another
// query or finishes.

#include <algorithm>
#include <chrono>
#include <cstring>
#include <map>

#include "kz_device.h"
#include "kz_kernels.h"

namespace kz {

struct QueryRef {
    const uint64_t* rowptr;
    const uint32_t* col;
};

__device__ __forceinline__ void flush_query(unsigned long long& acc, uint32_t q, unsigned long long* results) {
    const unsigned long long s = wave_sum_u64(acc);
    if ((threadIdx.x & 63) == 0 && s) atomicAdd(&results[q], s);
    acc = 0;
}

// ------------------------------------------------------------ wave tasks --
// One task per wave, |N+(u)| <= 64.  LDS per wave: L[64] ids, 128-slot hash,
// 64 bitmap rows (one u64 each).
template <int K>
__global__ __launch_bounds__(256) void clique_wave_kernel(const QueryRef* __restrict__ queries,
                                                          const uint32_t* __restrict__ task_v,
                                                          const uint32_t* __restrict__ task_q, uint32_t ntasks,
                                                          uint32_t* __restrict__ cursor,
                                                          unsigned long long* __restrict__ results) {
    constexpr int WPB = 4;
    constexpr uint32_t H = 128;
    __shared__ uint32_t sL[WPB][64];
    __shared__ uint32_t sKey[WPB][H];
    __shared__ uint16_t sVal[WPB][H];
    __shared__ unsigned long long sRow[WPB][64];
    const int lane = threadIdx.x & 63;
    const int w = threadIdx.x >> 6;
    uint32_t* L = sL[w];
    uint32_t* hk = sKey[w];
    uint16_t* hv = sVal[w];
    unsigned long long* rows = sRow[w];
    unsigned long long acc = 0;
    uint32_t qcur = 0xFFFFFFFFu;
    const uint64_t* rowptr = nullptr;
    const uint32_t* col = nullptr;
    while (true) {
        uint32_t t = 0;
        if (lane == 0) t = atomicAdd(cursor, 1u);
        t = __shfl(t, 0);
        if (t >= ntasks) break;
        const uint32_t q = task_q ? task_q[t] : 0u;
        if (q != qcur) {
            if (qcur != 0xFFFFFFFFu) flush_query(acc, qcur, results);
            qcur = q;
            rowptr = queries[q].rowptr;
            col = queries[q].col;
        }
        const uint32_t u = task_v[t];
        const uint64_t ub = rowptr[u];
        const uint32_t d = uint32_t(rowptr[u + 1] - ub);
        const uint32_t x = lane < d ? col[ub + lane] : 0xFFFFFFFFu;
        L[lane] = x;
        hk[lane] = kEmpty;
        hk[lane + 64] = kEmpty;
        rows[lane] = 0;
        wave_lds_sync();
        if (lane < d) hash_insert(hk, hv, H - 1, 25, x, uint16_t(lane));
        wave_lds_sync();
        const uint32_t lmax = L[d - 1];
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
    if (qcur != 0xFFFFFFFFu) flush_query(acc, qcur, results);
}

// ----------------------------------------------------------- block tasks --
// One task per workgroup.  Dynamic LDS: [rows Dmax*W u64 (k >= 4, rows in
// LDS)][L Dmax u32][hash keys H u32][hash vals H u16].  With GLOBAL_ROWS the
// rows live in an HBM slab of Dmax*W u64 per workgroup.
template <int K, bool GLOBAL_ROWS>
__global__ __launch_bounds__(512) void clique_block_kernel(const QueryRef* __restrict__ queries,
                                                           const uint32_t* __restrict__ task_v,
                                                           const uint32_t* __restrict__ task_q, uint32_t ntasks,
                                                           uint32_t* __restrict__ cursor,
                                                           unsigned long long* __restrict__ results, uint32_t Dmax,
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
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int nwaves = blockDim.x >> 6;
    const uint32_t hmask = H - 1;
    unsigned long long acc = 0;
    uint32_t qcur = 0xFFFFFFFFu;
    const uint64_t* rowptr = nullptr;
    const uint32_t* col = nullptr;
    while (true) {
        if (threadIdx.x == 0) sTask = atomicAdd(cursor, 1u);
        __syncthreads();
        const uint32_t t = sTask;
        if (t >= ntasks) break;
        const uint32_t q = task_q ? task_q[t] : 0u;
        if (q != qcur) {
            if (qcur != 0xFFFFFFFFu) flush_query(acc, qcur, results);
            qcur = q;
            rowptr = queries[q].rowptr;
            col = queries[q].col;
        }
        const uint32_t u = task_v[t];
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
                    const uint32_t qq = base + lane;
                    const uint32_t y = qq < dv ? col[vb + qq] : 0xFFFFFFFFu;
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
    if (qcur != 0xFFFFFFFFu) flush_query(acc, qcur, results);
}

__global__ void reset_u32_kernel(uint32_t* p, int n) {
    for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0;
}
__global__ void reset_u64_kernel(unsigned long long* p, int n) {
    for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0;
}

// ------------------------------------------------------------- host side --

BucketCfg default_buckets() {
    BucketCfg c;
    const uint32_t hi[kNumBuckets] = {64, 256, 512, 1024, 0xFFFFFFFFu};
    const uint32_t blk[kNumBuckets] = {256, 256, 256, 512, 512};
    for (int b = 0; b < kNumBuckets; ++b) {
        c.hi[b] = hi[b];
        c.block[b] = blk[b];
    }
    return c;
}

int check_buckets(const BucketCfg& c) {
    if (c.hi[0] < 2 || c.hi[0] > 64) KZ_FAIL(KZ_ERR_INVALID, "first bucket bound must be in 2..64");
    for (int b = 1; b < kNumBuckets - 1; ++b)
        if (c.hi[b] <= c.hi[b - 1]) KZ_FAIL(KZ_ERR_INVALID, "bucket bounds must increase");
    if (c.hi[kNumBuckets - 2] > 1024) KZ_FAIL(KZ_ERR_INVALID, "LDS buckets end at 1024 or below");
    for (int b = 0; b < kNumBuckets; ++b)
        if (c.block[b] != 256 && c.block[b] != 512) KZ_FAIL(KZ_ERR_INVALID, "block sizes must be 256 or 512");
    if (c.block[0] != 256) KZ_FAIL(KZ_ERR_INVALID, "the wave bucket runs 256-thread blocks");
    return KZ_OK;
}

static uint32_t bucket_of(const BucketCfg& c, uint32_t d) {
    for (int b = 0; b < kNumBuckets - 1; ++b)
        if (d <= c.hi[b]) return uint32_t(b);
    return kNumBuckets - 1;
}

CliqueJob::~CliqueJob() {
    if (exec) hipGraphExecDestroy(exec);
    if (graph) hipGraphDestroy(graph);
    hipFree(d_queries);
    hipFree(d_task_v);
    hipFree(d_task_q);
    hipFree(d_cursor);
    hipFree(d_result);
    for (auto& s : slabs) hipFree(s.second);
}

static size_t lds_bytes(uint32_t Dmax, uint32_t W, uint32_t H, bool rows_in_lds) {
    return (rows_in_lds ? size_t(Dmax) * W * 8 : 0) + size_t(Dmax) * 4 + size_t(H) * 4 + size_t(H) * 2 + 16;
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

// One (k, bucket) launch.
template <int K>
static int launch_group(CliqueJob& x, const CliqueJob::Group& gr, hipStream_t s) {
    const Runtime& R = rt();
    const uint32_t nt = uint32_t(gr.count);
    const QueryRef* qrefs = static_cast<const QueryRef*>(x.d_queries);
    const uint32_t* tv = x.d_task_v + gr.first;
    const uint32_t* tq = x.d_task_q ? x.d_task_q + gr.first : nullptr;
    uint32_t* cur = x.d_cursor + gr.cursor;
    if (gr.bucket == 0) {
        const uint32_t grid = std::min<uint32_t>((nt + 3) / 4, uint32_t(R.num_cus) * 8);
        hipLaunchKernelGGL(clique_wave_kernel<K>, dim3(grid), dim3(256), 0, s, qrefs, tv, tq, nt, cur,
                           x.d_result);
        return KZ_OK;
    }
    const uint32_t Dmax = gr.dmax;
    const uint32_t W = (Dmax + 63) / 64;
    uint32_t H = 64;
    while (H < 2 * Dmax) H <<= 1;
    uint32_t hbits = 0;
    while ((1u << hbits) < H) ++hbits;
    const uint32_t hshift = 32 - hbits;
    // only k >= 4 builds bitmap rows; the last bucket keeps them in HBM
    const bool global_rows = (gr.bucket == kNumBuckets - 1) && K >= 4;
    const size_t lds = lds_bytes(Dmax, W, H, !global_rows && K >= 4);
    const uint32_t block = x.cfg.block[gr.bucket];
    if (lds + kStaticLdsReserve > R.lds_per_block)
        KZ_FAIL(KZ_ERR_UNSUPPORTED, "bucket %d needs %zu bytes of LDS (limit %zu)", gr.bucket, lds, R.lds_per_block);
    uint32_t per_cu = std::max<uint32_t>(1, uint32_t(R.lds_per_block / (lds + kStaticLdsReserve)));
    per_cu = std::min<uint32_t>(per_cu, 2048 / block);
    uint32_t grid = std::min<uint32_t>(nt, uint32_t(R.num_cus) * per_cu);
    unsigned long long* slab = nullptr;
    if (global_rows) {
        auto it = x.slabs.find(K);
        if (it == x.slabs.end()) KZ_FAIL(KZ_ERR_STATE, "missing row slab for k=%d", K);
        slab = it->second;
        grid = std::min<uint32_t>(grid, x.slab_blocks[K]);
    }
    if (grid == 0) KZ_FAIL(KZ_ERR_STATE, "empty launch for bucket %d", gr.bucket);
    if (global_rows) {
        hipLaunchKernelGGL((clique_block_kernel<K, true>), dim3(grid), dim3(block), lds, s, qrefs, tv, tq, nt,
                           cur, x.d_result, Dmax, W, H, hshift, slab);
    } else {
        hipLaunchKernelGGL((clique_block_kernel<K, false>), dim3(grid), dim3(block), lds, s, qrefs, tv, tq, nt,
                           cur, x.d_result, Dmax, W, H, hshift, nullptr);
    }
    return KZ_OK;
}

// Builds the task groups of every (k, bucket), uploads them and captures one
// hipGraph: reset cursors and results, one parallel branch per group, join.
int clique_job_prepare(CliqueJob& x, const std::vector<CliqueQuery>& queries, const BucketCfg& cfg) {
    if (x.ready) return KZ_OK;
    if (int e = check_buckets(cfg)) return e;
    if (int e = set_lds_attrs()) return e;
    x.cfg = cfg;
    x.nqueries = uint32_t(queries.size());
    // (k, bucket) -> list of (degree, vertex, query)
    struct T {
        uint32_t d, v, q;
    };
    std::map<std::pair<int, int>, std::vector<T>> lists;
    for (uint32_t qi = 0; qi < queries.size(); ++qi) {
        const CliqueQuery& Q = queries[qi];
        if (Q.k < 3 || Q.k > 5) KZ_FAIL(KZ_ERR_UNSUPPORTED, "clique size %d outside 3..5", Q.k);
        const std::vector<uint32_t>& fd = *Q.fdeg;
        for (uint32_t v = 0; v < fd.size(); ++v) {
            if (fd[v] < uint32_t(Q.k - 1)) continue;
            if (fd[v] > kMaxBitmapDegree)
                KZ_FAIL(KZ_ERR_UNSUPPORTED,
                        "persistent clique form needs out-degree <= %u (found %u); orient the graph by degree",
                        kMaxBitmapDegree, fd[v]);
            lists[{Q.k, int(bucket_of(cfg, fd[v]))}].push_back(T{fd[v], v, qi});
        }
    }
    std::vector<uint32_t> hv, hq;
    x.groups.clear();
    x.ntasks = 0;
    int cursor = 0;
    for (auto& kv : lists) {
        auto& L = kv.second;
        std::stable_sort(L.begin(), L.end(), [](const T& a, const T& b) { return a.d > b.d; });
        CliqueJob::Group gr;
        gr.k = kv.first.first;
        gr.bucket = kv.first.second;
        gr.first = hv.size();
        gr.count = L.size();
        gr.dmax = gr.bucket == 0 ? 64 : (gr.bucket == kNumBuckets - 1 ? L.front().d : cfg.hi[gr.bucket]);
        gr.cursor = cursor++;
        for (const T& t : L) {
            hv.push_back(t.v);
            hq.push_back(t.q);
        }
        x.groups.push_back(gr);
        if (gr.bucket == kNumBuckets - 1 && gr.k >= 4) {
            const uint32_t W = (gr.dmax + 63) / 64;
            const uint32_t blocks = uint32_t(std::min<uint64_t>(gr.count, uint64_t(rt().num_cus) * 2));
            unsigned long long* slab = nullptr;
            KZ_HIP(hipMalloc(&slab, size_t(blocks) * gr.dmax * W * 8));
            x.slabs[gr.k] = slab;
            x.slab_blocks[gr.k] = blocks;
        }
    }
    x.ntasks = hv.size();
    std::vector<QueryRef> qr(std::max<size_t>(1, queries.size()));
    for (size_t i = 0; i < queries.size(); ++i) qr[i] = QueryRef{queries[i].rowptr, queries[i].col};
    KZ_HIP(hipMalloc(&x.d_queries, sizeof(QueryRef) * qr.size()));
    KZ_HIP(hipMemcpy(x.d_queries, qr.data(), sizeof(QueryRef) * qr.size(), hipMemcpyHostToDevice));
    KZ_HIP(hipMalloc(&x.d_task_v, sizeof(uint32_t) * std::max<size_t>(1, hv.size())));
    if (!hv.empty()) KZ_HIP(hipMemcpy(x.d_task_v, hv.data(), sizeof(uint32_t) * hv.size(), hipMemcpyHostToDevice));
    if (queries.size() > 1) {
        KZ_HIP(hipMalloc(&x.d_task_q, sizeof(uint32_t) * std::max<size_t>(1, hq.size())));
        if (!hq.empty()) KZ_HIP(hipMemcpy(x.d_task_q, hq.data(), sizeof(uint32_t) * hq.size(), hipMemcpyHostToDevice));
    }
    KZ_HIP(hipMalloc(&x.d_cursor, sizeof(uint32_t) * std::max(1, cursor)));
    KZ_HIP(hipMalloc(&x.d_result, sizeof(unsigned long long) * std::max<size_t>(1, queries.size())));
    // capture
    Runtime& R = rt();
    hipStream_t main = R.stream;
    const int nbr = std::max<int>(1, int(x.groups.size()));
    std::vector<hipStream_t> br(static_cast<size_t>(nbr), nullptr);
    std::vector<hipEvent_t> join(static_cast<size_t>(nbr), nullptr);
    hipEvent_t fork;
    for (int i = 0; i < nbr; ++i) {
        KZ_HIP(hipStreamCreateWithFlags(&br[i], hipStreamNonBlocking));
        KZ_HIP(hipEventCreateWithFlags(&join[i], hipEventDisableTiming));
    }
    KZ_HIP(hipEventCreateWithFlags(&fork, hipEventDisableTiming));
    KZ_HIP(hipStreamBeginCapture(main, hipStreamCaptureModeThreadLocal));
    hipLaunchKernelGGL(reset_u32_kernel, dim3(1), dim3(256), 0, main, x.d_cursor, std::max(1, cursor));
    hipLaunchKernelGGL(reset_u64_kernel, dim3(1), dim3(256), 0, main, x.d_result, int(std::max<size_t>(1, queries.size())));
    KZ_HIP(hipEventRecord(fork, main));
    int e = KZ_OK;
    for (size_t i = 0; i < x.groups.size() && e == KZ_OK; ++i) {
        KZ_HIP(hipStreamWaitEvent(br[i], fork, 0));
        const CliqueJob::Group& gr = x.groups[i];
        if (gr.k == 3) e = launch_group<3>(x, gr, br[i]);
        else if (gr.k == 4) e = launch_group<4>(x, gr, br[i]);
        else e = launch_group<5>(x, gr, br[i]);
    }
    for (size_t i = 0; i < x.groups.size(); ++i) {
        if (i >= x.groups.size()) break;
        KZ_HIP(hipEventRecord(join[i], br[i]));
        KZ_HIP(hipStreamWaitEvent(main, join[i], 0));
    }
    hipGraph_t graph;
    KZ_HIP(hipStreamEndCapture(main, &graph));
    if (e == KZ_OK && hipGetLastError() != hipSuccess) {
        e = KZ_ERR_DEVICE;
        kz::set_error("a bucket launch was rejected during capture");
    }
    for (int i = 0; i < nbr; ++i) {
        hipStreamDestroy(br[i]);
        hipEventDestroy(join[i]);
    }
    hipEventDestroy(fork);
    if (e) {
        hipGraphDestroy(graph);
        return e;
    }
    x.graph = graph;
    KZ_HIP(hipGraphInstantiate(&x.exec, graph, nullptr, nullptr, 0));
    x.ready = true;
    return KZ_OK;
}

int clique_job_run(CliqueJob& x, uint64_t* counts, double* device_ms) {
    Runtime& R = rt();
    hipEvent_t a, b;
    KZ_HIP(hipEventCreate(&a));
    KZ_HIP(hipEventCreate(&b));
    KZ_HIP(hipEventRecord(a, R.stream));
    KZ_HIP(hipGraphLaunch(x.exec, R.stream));
    KZ_HIP(hipEventRecord(b, R.stream));
    std::vector<unsigned long long> res(std::max<uint32_t>(1, x.nqueries));
    KZ_HIP(hipMemcpyAsync(res.data(), x.d_result, sizeof(unsigned long long) * res.size(), hipMemcpyDeviceToHost,
                          R.stream));
    KZ_HIP(hipStreamSynchronize(R.stream));
    float ms = 0;
    KZ_HIP(hipEventElapsedTime(&ms, a, b));
    hipEventDestroy(a);
    hipEventDestroy(b);
    for (uint32_t i = 0; i < x.nqueries; ++i) counts[i] = res[i];
    if (device_ms) *device_ms = ms;
    return KZ_OK;
}

// Candidate bucket configurations tried by the autotuner.
std::vector<BucketCfg> tuning_candidates() {
    std::vector<BucketCfg> out;
    const uint32_t w0s[] = {32, 64};
    const uint32_t b1s[] = {128, 256};
    const uint32_t blks[] = {256, 512};
    for (uint32_t w0 : w0s)
        for (uint32_t b1 : b1s)
            for (uint32_t bk : blks) {
                BucketCfg c = default_buckets();
                c.hi[0] = w0;
                c.hi[1] = b1;
                c.block[1] = c.block[2] = bk;
                out.push_back(c);
            }
    return out;
}

}  // namespace kz
