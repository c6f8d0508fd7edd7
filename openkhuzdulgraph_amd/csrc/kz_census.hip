// Induced 3- and 4-vertex motif census of a resident symmetric graph.
//
// Reference: the 3-vertex part is what the three_motif apps count
// (kautomine/three_motif.cc:59-139: induced wedges and triangles, divided
// by the multiplicities of :44).  The reference has no 4-vertex census app;
// the 4-vertex part follows the standard combinatorial route.  Non-induced
// counts come from degrees d, per-edge triangle counts t(e), 4-cycles and
// 4-cliques:
//   wedges = sum C(d,2)        3-stars  = sum C(d,3)
//   3-paths = sum_e (d_u-1)(d_v-1) - 3T
//   tailed triangles = sum_v t(v)(d_v-2) = 1/2 sum_e t(e)(d_u+d_v-4)
//   diamonds = sum_e C(t(e),2)
// and the induced counts by back-substitution through the containment
// matrix of the six connected 4-vertex graphs (host side, kz_motif_census).
//
// Pipeline, all on the GPU with the graph resident:
//   0. once per graph: relabel by rank = (degree, id) with two 64-bit radix
//      sorts, (degree << 32 | id) and (row << 32 | new id); the forward
//      part F (v > u) is the suffix of each relabelled row.  Degree ranking
//      keeps every forward list short (|N+(u)| = O(sqrt(m))).
//   1. triangles per forward edge: per root u, N+(u) is staged in LDS with
//      a hash and the rows N+(v), v in N+(u), are streamed as 16-byte groups
//      (kz_stream.h); a hit (v, w) closes triangle (u, v, w) and counts it
//      on the root's edges (u, v), (u, w) in LDS and on (v, w) with one
//      global atomic.
//   2. one pass over vertices and F edges for the degree / t(e) sums.
//   3. 4-cycles, each counted once from its highest-ranked vertex u: wedges
//      u - v - w with v, w ranked below u (prefixes of the sorted relabelled
//      lists) increment c(w) in a per-workgroup dense counter array in HBM,
//      which also records the touched w; then sum C(c(w), 2) and clear.
//   4. 4-cliques: the persistent clique kernels on F.

#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <memory>

#include "kz_device.h"
#include "kz_kernels.h"
#include "kz_stream.h"

namespace kz {

namespace {

constexpr uint32_t kSmallRoot = 64;  // forward degree handled by one wave

struct U64Cast {
    __host__ __device__ uint64_t operator()(uint32_t x) const { return x; }
};

// Census state kept with the graph.
struct Ranked {
    uint32_t n = 0;
    uint64_t nnz = 0, fnnz = 0;
    uint64_t* rowptr = nullptr;  // relabelled symmetric lists, sorted
    uint32_t* col = nullptr;
    uint32_t* deg = nullptr;  // degree by rank (n + 1 entries, last 0)
    uint64_t* frowptr = nullptr;  // forward part
    uint32_t* fcol = nullptr;
    std::vector<uint32_t> h_fdeg;
    uint32_t* tasks = nullptr;  // roots with d+ >= 2: small ones, then large ones
    uint32_t nsmall = 0, nlarge = 0, large_dmax = 0;
    uint32_t* troot = nullptr;  // per forward edge (u, v): triangles with u lowest
    uint32_t* tmid = nullptr;   // per forward edge: triangles with it as (v, w)
    uint32_t c4_grid = 0;
    uint32_t* cnt = nullptr;      // c4_grid * n counters
    uint32_t* touched = nullptr;  // c4_grid * n touched lists
    uint32_t* cursor = nullptr;   // 4 task cursors, kCursorStride words apart
    unsigned long long* sums = nullptr;  // 8 accumulators
    uint64_t bytes = 0;           // device memory held (with the 4-clique job)
    CliqueJob k4;
    ~Ranked() {
        for (void* p : {(void*)rowptr, (void*)col, (void*)deg, (void*)frowptr, (void*)fcol, (void*)tasks, (void*)troot,
                        (void*)tmid, (void*)cnt, (void*)touched, (void*)cursor, (void*)sums})
            if (p) (void)hipFree(p);
    }
};

__global__ void rank_keys_kernel(const uint32_t* __restrict__ deg, uint32_t n, uint64_t* __restrict__ keys) {
    const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v < n) keys[v] = (uint64_t(deg[v]) << 32) | v;
}

// sorted[i] = (deg << 32 | v) in rank order: rank[v] = i, ndeg[i] = deg.
__global__ void rank_scatter_kernel(const uint64_t* __restrict__ sorted, uint32_t n, uint32_t* __restrict__ rank,
                                    uint32_t* __restrict__ order, uint32_t* __restrict__ ndeg) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i > n) return;
    if (i == n) {
        ndeg[n] = 0;
        return;
    }
    const uint64_t k = sorted[i];
    const uint32_t v = uint32_t(k);
    rank[v] = i;
    order[i] = v;
    ndeg[i] = uint32_t(k >> 32);
}

// Wave per new vertex i: its relabelled list as (i << 32 | rank[w]) keys.
__global__ void relabel_keys_kernel(const uint64_t* __restrict__ rowptr, const uint32_t* __restrict__ col,
                                    const uint32_t* __restrict__ order, const uint32_t* __restrict__ rank,
                                    const uint64_t* __restrict__ nrowptr, uint32_t n, uint64_t* __restrict__ keys) {
    const uint32_t i = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const uint32_t lane = threadIdx.x & 63;
    if (i >= n) return;
    const uint32_t v = order[i];
    const uint64_t src = rowptr[v];
    const uint64_t dst = nrowptr[i];
    const uint32_t d = uint32_t(nrowptr[i + 1] - dst);
    for (uint32_t k = lane; k < d; k += 64) keys[dst + k] = (uint64_t(i) << 32) | rank[col[src + k]];
}

__global__ void low_word_kernel(const uint64_t* __restrict__ keys, uint64_t m, uint32_t* __restrict__ out) {
    const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < m) out[i] = uint32_t(keys[i]);
}

__global__ void fill_u32_kernel(uint32_t* __restrict__ p, uint64_t m) {
    for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < m; i += uint64_t(gridDim.x) * blockDim.x)
        p[i] = 0;
}

__global__ void fill_small_kernel(uint32_t* __restrict__ a, int na, unsigned long long* __restrict__ b, int nb) {
    for (int i = threadIdx.x; i < na; i += blockDim.x) a[i] = 0;
    for (int i = threadIdx.x; i < nb; i += blockDim.x) b[i] = 0;
}

// Pass 1: triangles per forward edge.  One root per workgroup of T threads
// (T = 64: roots with d+ <= 64, one wave).  Dynamic LDS: [row descriptors
// Dmax u64][hash keys H u32][L Dmax u32][edge counters Dmax u32][group
// prefix Dmax+1 u32][filter FW u32][hash vals H u16].
template <int T>
__global__ __launch_bounds__(T) void census_tri_kernel(const uint64_t* __restrict__ frowptr,
                                                       const uint32_t* __restrict__ fcol,
                                                       const uint32_t* __restrict__ tasks, uint32_t ntasks,
                                                       uint32_t* __restrict__ cursor, uint32_t* __restrict__ troot,
                                                       uint32_t* __restrict__ tmid, uint32_t Dmax, uint32_t H,
                                                       uint32_t hshift, uint32_t FW, uint32_t fshift) {
    extern __shared__ __align__(16) unsigned char smem[];
    unsigned char* p = smem;
    unsigned long long* rb = reinterpret_cast<unsigned long long*>(p);
    p += size_t(Dmax) * 8;
    uint32_t* hk = reinterpret_cast<uint32_t*>(p);  // 8-byte aligned (pair reads)
    p += size_t(H) * 4;
    uint32_t* L = reinterpret_cast<uint32_t*>(p);
    p += size_t(Dmax) * 4;
    uint32_t* cnt = reinterpret_cast<uint32_t*>(p);
    p += size_t(Dmax) * 4;
    uint32_t* rp = reinterpret_cast<uint32_t*>(p);
    p += size_t(Dmax + 1) * 4;
    uint32_t* fb = reinterpret_cast<uint32_t*>(p);
    p += size_t(FW) * 4;
    uint16_t* hv = reinterpret_cast<uint16_t*>(p);
    __shared__ uint32_t sTask, sTmp[T / 64], sSpan;
    const RootSet rs{hk, hv, H - 1, hshift, fb, fshift};
    TaskChunk chunk(ntasks, gridDim.x, T == 64 ? 2 : 8);
    while (true) {
        if (threadIdx.x == 0) sTask = thread_next_task(chunk, cursor);
        __syncthreads();
        const uint32_t t = wave_bcast(sTask);
        if (t >= ntasks) break;
        const uint32_t u = tasks[t];
        const uint64_t ub = ldg(frowptr, u);
        const uint32_t d = uint32_t(ldg(frowptr, u + 1) - ub);
        for (uint32_t i = threadIdx.x; i < d; i += T) {
            L[i] = ldg(fcol, ub + i);
            cnt[i] = 0;
        }
        root_clear(rs, H, FW, threadIdx.x, T);
        __syncthreads();
        for (uint32_t i = threadIdx.x; i < d; i += T) root_insert(rs, L[i], uint16_t(i));
        const uint32_t nr = d - 1;
        if (threadIdx.x == 0) sSpan = 0;
        for (uint32_t r = threadIdx.x; r <= nr; r += T) {
            uint32_t groups = 0, len = 0;
            uint64_t desc = 0;
            if (r < nr) {
                row_desc(frowptr, L[r], d - 1 - r, &desc, &groups, &len);
                rb[r] = desc;
            }
            rp[r] = groups;
        }
        __syncthreads();
        const uint32_t G = block_exclusive_scan(rp, nr + 1, sTmp);
        auto hit = [&](uint32_t r, uint32_t j, uint64_t at) {
            atomicAdd(&cnt[r], 1u);
            atomicAdd(&cnt[j], 1u);
            atomicAdd(&tmid[at], 1u);
        };
        stream_groups(fcol, rb, rp, nr, G, threadIdx.x, T, rs, hit, T > 64 ? &sSpan : nullptr);
        // skewed rows: each wave searches the candidates of its own rows
        search_skewed(fcol, rb, L, d, 0, nr, threadIdx.x >> 6, T / 64, hit);
        __syncthreads();
        for (uint32_t i = threadIdx.x; i < d; i += T) troot[ub + i] = cnt[i];
        __syncthreads();
    }
}

// Pass 2: sums[0..6) += {C(d,2), C(d,3), (d_u-1)(d_v-1), t(e), C(t(e),2),
// t(e)(d_u+d_v-4)} over vertices u and forward edges e = (u, v).
__global__ __launch_bounds__(256) void census_sums_kernel(const uint64_t* __restrict__ frowptr,
                                                          const uint32_t* __restrict__ fcol,
                                                          const uint32_t* __restrict__ deg, uint32_t n,
                                                          const uint32_t* __restrict__ troot,
                                                          const uint32_t* __restrict__ tmid,
                                                          unsigned long long* __restrict__ sums) {
    unsigned long long s[6] = {0, 0, 0, 0, 0, 0};
    for (uint32_t u = blockIdx.x * blockDim.x + threadIdx.x; u < n; u += gridDim.x * blockDim.x) {
        const uint64_t du = deg[u];
        s[0] += du * (du - 1) / 2;
        s[1] += du * (du - 1) * (du - 2) / 6;
        for (uint64_t e = frowptr[u]; e < frowptr[u + 1]; ++e) {
            const uint64_t dv = deg[fcol[e]];
            const uint64_t t = uint64_t(troot[e]) + tmid[e];
            s[2] += (du - 1) * (dv - 1);
            s[3] += t;
            s[4] += t * (t - 1) / 2;
            s[5] += t * (du + dv - 4);  // t > 0 implies du, dv >= 2
        }
    }
#pragma unroll
    for (int i = 0; i < 6; ++i) {
        const unsigned long long x = wave_sum_u64(s[i]);
        if ((threadIdx.x & 63) == 0 && x) atomicAdd(&sums[i], x);
    }
}

// Pass 3: 4-cycles.  Tasks are all vertices, highest rank first.  A root
// whose wedges (bounded by the degrees of its neighbours ranked below it)
// fit an LDS hash of kC4Slots counters counts there; the others use the
// workgroup's n counters and n-entry touched list in HBM (one returning
// atomic per wedge, the cost of this pass before the LDS counters).
constexpr uint32_t kC4Slots = 8192;  // 64 KB of LDS (4096: 266.4 ms per census, 8192: 263.7)
constexpr int kC4Block = 1024;

template <int T>
__global__ __launch_bounds__(T) void census_c4_kernel(const uint64_t* __restrict__ rowptr,
                                                        const uint32_t* __restrict__ col, uint32_t n,
                                                        uint32_t* __restrict__ cursor, uint32_t* __restrict__ cnt_slab,
                                                        uint32_t* __restrict__ touched_slab,
                                                        unsigned long long* __restrict__ sums) {
    __shared__ uint32_t sTask, sTouched, sBound;
    __shared__ uint32_t sKey[kC4Slots], sCnt[kC4Slots];
    uint32_t* cnt = cnt_slab + size_t(blockIdx.x) * n;
    uint32_t* touched = touched_slab + size_t(blockIdx.x) * n;
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t wave = threadIdx.x >> 6;
    const uint32_t nw = blockDim.x >> 6;
    unsigned long long acc = 0;
    if (threadIdx.x == 0) sTouched = 0;
    TaskChunk chunk(n, gridDim.x, 8);
    while (true) {
        if (threadIdx.x == 0) {
            sTask = thread_next_task(chunk, cursor);
            sBound = 0;
        }
        __syncthreads();
        const uint32_t t = wave_bcast(sTask);
        if (t >= n) break;
        const uint32_t u = n - 1 - t;
        const uint64_t ub = ldg(rowptr, u);
        const uint32_t du = uint32_t(ldg(rowptr, u + 1) - ub);
        const uint32_t pre = lower_bound_global(col + ub, du, u);  // neighbours ranked below u
        // wedge bound: the degrees of the neighbours below u
        // (64-bit sums, each wave's clamped to 2^16 before the LDS add, so
        // that no sum wraps and a root above the LDS size never takes it)
        unsigned long long part = 0;
        for (uint32_t k = threadIdx.x; k < pre; k += blockDim.x) {
            const uint32_t v = ldg(col, ub + k);
            part += ldg(rowptr, v + 1) - ldg(rowptr, v);
        }
        part = wave_sum_u64(part);
        if (lane == 0 && part) atomicAdd(&sBound, uint32_t(part < 65536ull ? part : 65536ull));
        __syncthreads();
        const uint32_t bound = wave_bcast(sBound);
        if (bound <= kC4Slots / 2) {
            // LDS counters: an open-addressing hash of at least twice the
            // wedges, so every probe ends
            uint32_t H = 64;
            while (H < 2 * bound) H <<= 1;
            uint32_t bits = 0;
            while ((1u << bits) < H) ++bits;
            const uint32_t hshift = 32 - bits;
            for (uint32_t i = threadIdx.x; i < H; i += blockDim.x) {
                sKey[i] = kEmpty;
                sCnt[i] = 0;
            }
            __syncthreads();
            for (uint32_t k = wave; k < pre; k += nw) {
                const uint32_t v = ldg(col, ub + k);
                const uint64_t vb = ldg(rowptr, v);
                const uint32_t dv = uint32_t(ldg(rowptr, v + 1) - vb);
                for (uint32_t base = 0; base < dv; base += 64) {
                    const uint32_t i = base + lane;
                    const uint32_t w = i < dv ? ldg(col, vb + i) : 0xFFFFFFFFu;
                    if (w < u) {
                        uint32_t h = hash_slot(w, hshift);
                        while (true) {
                            const uint32_t key = sKey[h];
                            if (key == w) break;
                            if (key == kEmpty) {
                                const uint32_t old = atomicCAS(&sKey[h], kEmpty, w);
                                if (old == kEmpty || old == w) break;
                            }
                            h = (h + 1) & (H - 1);
                        }
                        atomicAdd(&sCnt[h], 1u);
                    }
                    if (__ballot(w >= u)) break;  // sorted lists: the rest ranks above u
                }
            }
            __syncthreads();
            for (uint32_t i = threadIdx.x; i < H; i += blockDim.x) {
                const uint64_t c = sCnt[i];
                acc += c * (c - 1) / 2;
            }
            __syncthreads();
            continue;
        }
        for (uint32_t k = wave; k < pre; k += nw) {
            const uint32_t v = ldg(col, ub + k);
            const uint64_t vb = ldg(rowptr, v);
            const uint32_t dv = uint32_t(ldg(rowptr, v + 1) - vb);
            for (uint32_t base = 0; base < dv; base += 64) {
                const uint32_t i = base + lane;
                const uint32_t w = i < dv ? ldg(col, vb + i) : 0xFFFFFFFFu;
                if (w < u && atomicAdd(&cnt[w], 1u) == 0) touched[atomicAdd(&sTouched, 1u)] = w;
                if (__ballot(w >= u)) break;  // sorted lists: the rest ranks above u
            }
        }
        __syncthreads();
        const uint32_t nt = sTouched;
        for (uint32_t s = threadIdx.x; s < nt; s += blockDim.x) {
            const uint32_t w = touched[s];
            const uint64_t c = cnt[w];
            acc += c * (c - 1) / 2;
            cnt[w] = 0;
        }
        __syncthreads();
        if (threadIdx.x == 0) sTouched = 0;
    }
    const unsigned long long x = wave_sum_u64(acc);
    if (lane == 0 && x) atomicAdd(&sums[6], x);
}

static void hash_geometry(uint32_t Dmax, uint32_t* H, uint32_t* hshift) {
    uint32_t h = 64;
    while (h < 2 * Dmax) h <<= 1;
    uint32_t bits = 0;
    while ((1u << bits) < h) ++bits;
    *H = h;
    *hshift = 32 - bits;
}

static size_t tri_lds_bytes(uint32_t Dmax, uint32_t H) {
    uint32_t fw, fs;
    filter_geometry(Dmax, &fw, &fs);
    return size_t(Dmax) * 8 + size_t(Dmax) * 8 + size_t(Dmax + 1) * 4 + size_t(fw) * 4 + size_t(H) * 6 + 16;
}

static int sort_u64(uint64_t* in, uint64_t* out, uint64_t m, int end_bit, Scratch& tmp, hipStream_t s) {
    size_t bytes = 0;
    KZ_HIP(hipcub::DeviceRadixSort::SortKeys(nullptr, bytes, in, out, int(m), 0, end_bit, s));
    if (int e = scratch_reserve(tmp, bytes)) return e;
    KZ_HIP(hipcub::DeviceRadixSort::SortKeys(tmp.ptr, bytes, in, out, int(m), 0, end_bit, s));
    return KZ_OK;
}

static int scan_to_u64(const uint32_t* in, uint64_t* out, uint64_t n1, Scratch& tmp, hipStream_t s) {
    hipcub::TransformInputIterator<uint64_t, U64Cast, const uint32_t*> it(in, U64Cast());
    size_t bytes = 0;
    KZ_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, it, out, n1, s));
    if (int e = scratch_reserve(tmp, bytes)) return e;
    KZ_HIP(hipcub::DeviceScan::ExclusiveSum(tmp.ptr, bytes, it, out, n1, s));
    return KZ_OK;
}

static int bits_for(uint64_t x) {
    int b = 0;
    while (b < 64 && (uint64_t(1) << b) <= x) ++b;
    return b;
}

// Step 0: relabel by (degree, id) rank and split off the forward part.
static int build_ranked(DeviceGraph& g, Ranked& R) {
    const Runtime& rt_ = rt();
    hipStream_t s = rt_.stream;
    const uint32_t n = g.n;
    if (g.nnz >= (uint64_t(1) << 31)) KZ_FAIL(KZ_ERR_UNSUPPORTED, "census relabelling supports < 2^31 list entries");
    R.n = n;
    R.nnz = g.nnz;
    R.bytes = 4 * (size_t(n) + 1) + 8 * (size_t(n) + 1) + col_alloc_bytes(g.nnz) + 8 * (size_t(n) + 1);
    Scratch tmp;
    uint64_t *k0 = nullptr, *k1 = nullptr;
    uint32_t *rank = nullptr, *order = nullptr, *skip = nullptr, *fdeg = nullptr;
    // the temporaries are freed on every return, failures included
    auto cleanup = [&]() {
        for (void** p : {(void**)&k0, (void**)&k1, (void**)&rank, (void**)&order, (void**)&skip, (void**)&fdeg,
                         &tmp.ptr})
            if (*p) {
                (void)hipFree(*p);
                *p = nullptr;
            }
    };
    struct Cleanup {
        decltype(cleanup)& f;
        ~Cleanup() { f(); }
    } cleanup_on_return{cleanup};
    auto fail = [&](int e) { return e; };
    const uint64_t kmax = std::max<uint64_t>(n, g.nnz);
    if (hipMalloc(&k0, 8 * std::max<uint64_t>(1, kmax)) != hipSuccess ||
        hipMalloc(&k1, 8 * std::max<uint64_t>(1, kmax)) != hipSuccess ||
        hipMalloc(&rank, 4 * (size_t(n) + 1)) != hipSuccess || hipMalloc(&order, 4 * (size_t(n) + 1)) != hipSuccess ||
        hipMalloc(&R.deg, 4 * (size_t(n) + 1)) != hipSuccess ||
        hipMalloc(&R.rowptr, 8 * (size_t(n) + 1)) != hipSuccess ||
        hipMalloc(&R.col, col_alloc_bytes(g.nnz)) != hipSuccess) {
        (void)hipGetLastError();
        KZ_FAIL(KZ_ERR_NOMEM, "census relabelling buffers (%llu entries)", (unsigned long long)g.nnz);
    }
    const uint32_t nb = (n + 255) / 256;
    if (n) {
        hipLaunchKernelGGL(rank_keys_kernel, dim3(nb), dim3(256), 0, s, g.deg, n, k0);
        if (int e = sort_u64(k0, k1, n, 64, tmp, s)) return fail(e);
    }
    hipLaunchKernelGGL(rank_scatter_kernel, dim3((n + 1 + 255) / 256), dim3(256), 0, s, k1, n, rank, order, R.deg);
    if (int e = scan_to_u64(R.deg, R.rowptr, uint64_t(n) + 1, tmp, s)) return fail(e);
    if (g.nnz) {
        hipLaunchKernelGGL(relabel_keys_kernel, dim3((uint64_t(n) * 64 + 255) / 256), dim3(256), 0, s, g.rowptr,
                           g.col, order, rank, R.rowptr, n, k0);
        if (int e = sort_u64(k0, k1, g.nnz, 32 + bits_for(n), tmp, s)) return fail(e);
        hipLaunchKernelGGL(low_word_kernel, dim3((g.nnz + 255) / 256), dim3(256), 0, s, k1, g.nnz, R.col);
    }
    // forward part
    if (hipMalloc(&skip, 4 * (size_t(n) + 1)) != hipSuccess || hipMalloc(&fdeg, 4 * (size_t(n) + 1)) != hipSuccess ||
        hipMalloc(&R.frowptr, 8 * (size_t(n) + 1)) != hipSuccess) {
        (void)hipGetLastError();
        KZ_FAIL(KZ_ERR_NOMEM, "census forward split");
    }
    KZ_HIP(hipMemsetAsync(fdeg + n, 0, 4, s));
    if (n) hipLaunchKernelGGL(fwd_start_kernel, dim3(nb), dim3(256), 0, s, R.rowptr, R.col, n, skip, fdeg);
    if (int e = scan_to_u64(fdeg, R.frowptr, uint64_t(n) + 1, tmp, s)) return fail(e);
    R.h_fdeg.resize(n);
    KZ_HIP(hipMemcpyAsync(R.h_fdeg.data(), fdeg, 4 * size_t(n), hipMemcpyDeviceToHost, s));
    KZ_HIP(hipMemcpyAsync(&R.fnnz, R.frowptr + n, 8, hipMemcpyDeviceToHost, s));
    KZ_HIP(hipStreamSynchronize(s));
    R.bytes += col_alloc_bytes(R.fnnz);
    if (hipMalloc(&R.fcol, col_alloc_bytes(R.fnnz)) != hipSuccess) {
        (void)hipGetLastError();
        KZ_FAIL(KZ_ERR_NOMEM, "census forward lists");
    }
    if (n)
        hipLaunchKernelGGL(fwd_copy_kernel, dim3((uint64_t(n) * 64 + 255) / 256), dim3(256), 0, s, R.rowptr, R.col, n,
                           skip, R.frowptr, R.fcol);
    KZ_HIP(hipGetLastError());
    KZ_HIP(hipStreamSynchronize(s));
    return KZ_OK;
}

// Task lists, per-edge arrays, counter slabs and the 4-clique job.
static int prepare_passes(Ranked& R) {
    const Runtime& rt_ = rt();
    std::vector<uint32_t> small, large;
    for (uint32_t v = 0; v < R.n; ++v) {
        const uint32_t d = R.h_fdeg[v];
        if (d < 2) continue;
        if (d <= kSmallRoot) small.push_back(v);
        else {
            large.push_back(v);
            R.large_dmax = std::max(R.large_dmax, d);
        }
    }
    auto by_deg = [&](uint32_t a, uint32_t b) { return R.h_fdeg[a] > R.h_fdeg[b]; };
    std::stable_sort(small.begin(), small.end(), by_deg);
    std::stable_sort(large.begin(), large.end(), by_deg);
    R.nsmall = uint32_t(small.size());
    R.nlarge = uint32_t(large.size());
    std::vector<uint32_t> all(small);
    all.insert(all.end(), large.begin(), large.end());
    const size_t lds_limit = rt_.lds_per_block - kStaticLdsReserve;
    uint32_t H, hshift;
    hash_geometry(R.large_dmax, &H, &hshift);
    if (R.nlarge && tri_lds_bytes(R.large_dmax, H) > lds_limit)
        KZ_FAIL(KZ_ERR_UNSUPPORTED, "forward degree %u exceeds the census LDS budget", R.large_dmax);
    KZ_HIP(hipMalloc(&R.tasks, 4 * std::max<size_t>(1, all.size())));
    if (!all.empty()) KZ_HIP(hipMemcpy(R.tasks, all.data(), 4 * all.size(), hipMemcpyHostToDevice));
    KZ_HIP(hipMalloc(&R.troot, 4 * std::max<uint64_t>(1, R.fnnz)));
    KZ_HIP(hipMalloc(&R.tmid, 4 * std::max<uint64_t>(1, R.fnnz)));
    KZ_HIP(hipMalloc(&R.cursor, 4 * 4 * kCursorStride));
    KZ_HIP(hipMalloc(&R.sums, 8 * 8));
    // 4-cycle workgroups: two per CU, each with 8 bytes per vertex of slab,
    // fewer when the slabs would take more than a quarter of free memory
    size_t free_b = 0, total_b = 0;
    KZ_HIP(hipMemGetInfo(&free_b, &total_b));
    const size_t per_wg = size_t(std::max<uint32_t>(1, R.n)) * 8;
    R.c4_grid = uint32_t(std::max<size_t>(1, std::min<size_t>(size_t(rt_.num_cus) * 2, free_b / 4 / per_wg)));
    const size_t slab = size_t(R.c4_grid) * std::max<uint32_t>(1, R.n) * 4;
    R.bytes += 2 * slab + 8 * (2 * std::max<uint64_t>(1, R.fnnz) + all.size() + 8);
    if (hipMalloc(&R.cnt, slab) != hipSuccess || hipMalloc(&R.touched, slab) != hipSuccess) {
        (void)hipGetLastError();
        KZ_FAIL(KZ_ERR_NOMEM, "4-cycle counter slabs (%zu bytes each)", slab);
    }
    // counters start at zero and every pass leaves them at zero
    hipLaunchKernelGGL(fill_u32_kernel, dim3(1024), dim3(256), 0, rt_.stream, R.cnt, uint64_t(slab / 4));
    KZ_HIP(hipGetLastError());
    KZ_HIP(hipStreamSynchronize(rt_.stream));
    std::vector<CliqueQuery> q = {CliqueQuery{R.frowptr, R.fcol, &R.h_fdeg, 4}};
    if (int e = clique_job_prepare(R.k4, q, default_buckets())) return e;
    R.bytes += R.k4.bytes;
    return KZ_OK;
}

static int census_state(DeviceGraph& g, Ranked** out) {
    if (!g.census) {
        Ranked* R = new (std::nothrow) Ranked();
        if (!R) KZ_FAIL(KZ_ERR_NOMEM, "census state");
        std::shared_ptr<void> holder(R, [](void* p) { delete static_cast<Ranked*>(p); });
        if (int e = build_ranked(g, *R)) return e;
        if (int e = prepare_passes(*R)) return e;
        g.census = holder;
    }
    *out = static_cast<Ranked*>(g.census.get());
    return KZ_OK;
}

// Passes 1-4; sums[0..8) = {C(d,2), C(d,3), P, T3, DIA, PAW2, C4, K4}.
static int run_census(Ranked& R, uint64_t* sums, double* device_ms) {
    const Runtime& rt_ = rt();
    hipStream_t s = rt_.stream;
    hipEvent_t e0, e1;
    KZ_HIP(hipEventCreate(&e0));
    KZ_HIP(hipEventCreate(&e1));
    KZ_HIP(hipEventRecord(e0, s));
    hipLaunchKernelGGL(fill_small_kernel, dim3(1), dim3(64), 0, s, R.cursor, int(4 * kCursorStride), R.sums, 8);
    const uint64_t fz = std::max<uint64_t>(1, R.fnnz);
    hipLaunchKernelGGL(fill_u32_kernel, dim3(2048), dim3(256), 0, s, R.troot, fz);
    hipLaunchKernelGGL(fill_u32_kernel, dim3(2048), dim3(256), 0, s, R.tmid, fz);
    if (R.nsmall) {
        uint32_t H, hs;
        hash_geometry(kSmallRoot, &H, &hs);
        uint32_t fw, fs;
        filter_geometry(kSmallRoot, &fw, &fs);
        const uint32_t grid = std::min<uint32_t>(R.nsmall, uint32_t(rt_.num_cus) * 32);
        hipLaunchKernelGGL(census_tri_kernel<64>, dim3(grid), dim3(64), tri_lds_bytes(kSmallRoot, H), s, R.frowptr,
                           R.fcol, R.tasks, R.nsmall, R.cursor, R.troot, R.tmid, kSmallRoot, H, hs, fw, fs);
    }
    if (R.nlarge) {
        uint32_t H, hs;
        hash_geometry(R.large_dmax, &H, &hs);
        const size_t lds = tri_lds_bytes(R.large_dmax, H);
        // advisory on this runtime: a refusal only matters above 64 KiB
        if (hipFuncSetAttribute(reinterpret_cast<const void*>(&census_tri_kernel<256>),
                                hipFuncAttributeMaxDynamicSharedMemorySize,
                                int(rt_.lds_per_block - kStaticLdsReserve)) != hipSuccess)
            (void)hipGetLastError();
        uint32_t per_cu = std::max<uint32_t>(1, uint32_t(rt_.lds_per_block / (lds + kStaticLdsReserve)));
        per_cu = std::min<uint32_t>(per_cu, 8);
        const uint32_t grid = std::min<uint32_t>(R.nlarge, uint32_t(rt_.num_cus) * per_cu);
        uint32_t fw, fs;
        filter_geometry(R.large_dmax, &fw, &fs);
        hipLaunchKernelGGL(census_tri_kernel<256>, dim3(grid), dim3(256), lds, s, R.frowptr, R.fcol,
                           R.tasks + R.nsmall, R.nlarge, R.cursor + kCursorStride, R.troot, R.tmid, R.large_dmax, H, hs, fw, fs);
    }
    hipLaunchKernelGGL(census_sums_kernel, dim3(std::max(1, rt_.num_cus * 4)), dim3(256), 0, s, R.frowptr, R.fcol,
                       R.deg, R.n, R.troot, R.tmid, R.sums);
    if (R.n)
        // 16 waves per workgroup: the counter slabs bound the workgroups
        // (two per CU), and the HBM counters' returning atomics need many
        // waves in flight
        hipLaunchKernelGGL(census_c4_kernel<kC4Block>, dim3(R.c4_grid), dim3(kC4Block), 0, s, R.rowptr, R.col, R.n,
                           R.cursor + 2 * kCursorStride, R.cnt, R.touched, R.sums);
    KZ_HIP(hipGetLastError());
    KZ_HIP(hipEventRecord(e1, s));
    unsigned long long h[8] = {0};
    KZ_HIP(hipMemcpyAsync(h, R.sums, 8 * 8, hipMemcpyDeviceToHost, s));
    KZ_HIP(hipStreamSynchronize(s));
    float ms = 0;
    KZ_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    uint64_t k4 = 0;
    double k4_ms = 0;
    if (int e = clique_job_run(R.k4, &k4, &k4_ms)) return e;
    for (int i = 0; i < 7; ++i) sums[i] = h[i];
    sums[7] = k4;
    *device_ms = double(ms) + k4_ms;
    return KZ_OK;
}

}  // namespace

uint64_t census_bytes(const DeviceGraph& d) {
    return d.census ? static_cast<const Ranked*>(d.census.get())->bytes : 0;
}

int motif_census(DeviceGraph& d, uint64_t* counts, kz_run_stats* stats) {
    if (d.oriented) KZ_FAIL(KZ_ERR_INVALID, "the census needs a symmetric graph");
    auto t0 = std::chrono::steady_clock::now();
    Ranked* R = nullptr;
    if (int e = census_state(d, &R)) return e;
    uint64_t S[8];
    double ms = 0;
    if (int e = run_census(*R, S, &ms)) return e;
    // non-induced counts, then the induced ones by back-substitution through
    // the containment matrix (rows: copies of the graph in 3-path, 3-star,
    // 4-cycle, tailed triangle, diamond, 4-clique):
    //   3-path 1 0 4 2 6 12 | 3-star 0 1 0 1 2 4 | 4-cycle 0 0 1 0 1 3
    //   tailed 0 0 0 1 4 12 | diamond 0 0 0 0 1 6 | 4-clique 0 0 0 0 0 1
    const uint64_t T = S[3] / 3;
    const uint64_t nP3 = S[2] - 3 * T, nStar = S[1], nC4 = S[6], nPaw = S[5] / 2, nDia = S[4], nK4 = S[7];
    const uint64_t iK4 = nK4;
    const uint64_t iDia = nDia - 6 * iK4;
    const uint64_t iPaw = nPaw - 4 * iDia - 12 * iK4;
    const uint64_t iC4 = nC4 - iDia - 3 * iK4;
    const uint64_t iStar = nStar - iPaw - 2 * iDia - 4 * iK4;
    const uint64_t iP3 = nP3 - 4 * iC4 - 2 * iPaw - 6 * iDia - 12 * iK4;
    counts[0] = S[0] - 3 * T;
    counts[1] = T;
    counts[2] = iP3;
    counts[3] = iStar;
    counts[4] = iC4;
    counts[5] = iPaw;
    counts[6] = iDia;
    counts[7] = iK4;
    if (stats) {
        std::memset(stats, 0, sizeof(*stats));
        stats->device_ms = ms;
        stats->wall_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        stats->launches = 6;
        stats->tasks = uint64_t(R->nsmall) + R->nlarge + R->n;
        stats->embeddings[2] = uint64_t(d.nnz) / 2;
        stats->embeddings[3] = T;
        stats->embeddings[4] = iK4;
    }
    return KZ_OK;
}

}  // namespace kz

extern "C" int kz_motif_census(kz_graph g, uint64_t* counts, kz_run_stats* stats) {
    if (!g || !counts) KZ_FAIL(KZ_ERR_INVALID, "null argument");
    if (int e = kz::ensure_ready()) return e;
    return kz::motif_census(g->g, counts, stats);
}
