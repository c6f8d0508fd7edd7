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
#include <cstdlib>
#include <cstring>
#include <memory>

#include "kz_device.h"
#include "kz_kernels.h"
#include "kz_stream.h"

namespace kz {

namespace {

constexpr uint32_t kSmallRoot = 64;  // forward degree handled by one wave
constexpr int kCursors = 6;          // triangles (small, large roots), 4-cycles (wave, block, heavy)

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
    uint64_t* c4_bound = nullptr;  // per vertex: the 4-cycle pass's wedge bound b(u)
    uint32_t* c4_tasks = nullptr;  // roots of the wave, block and heavy 4-cycle kernels, in that order
    uint32_t n_c4_wave = 0, n_c4_block = 0, n_c4_heavy = 0;
    uint32_t c4_heavy_grid = 0;
    uint32_t c4_scratch_keys = 0;    // endpoints per heavy workgroup
    uint32_t c4_max_slots = 0;       // bucket hash slots used (a power of two <= kC4HeavySlots)
    uint32_t c4_direct_bound = 0;    // heavy roots up to this bound first try one unbucketed pass
    uint32_t* c4_scratch = nullptr;  // c4_heavy_grid * c4_scratch_keys endpoints
    uint32_t* cursor = nullptr;      // kCursors task cursors, kCursorStride words apart
    unsigned long long* sums = nullptr;  // 8 accumulators
    uint64_t bytes = 0;           // device memory held (with the 4-clique job)
    CliqueJob k4;
    ~Ranked() {
        for (void* p : {(void*)rowptr, (void*)col, (void*)deg, (void*)frowptr, (void*)fcol, (void*)tasks, (void*)troot,
                        (void*)tmid, (void*)c4_bound, (void*)c4_tasks, (void*)c4_scratch, (void*)cursor, (void*)sums})
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

// Pass 3: 4-cycles, each counted once from its highest-ranked vertex u as
// the pairs of wedges u - v - w, u - x - w (v, w, x ranked below u) with the
// same endpoint w: sum over w of C(c(w), 2), c(w) = the wedges ending at w.
// Roots are split by their wedge bound b(u) = sum of d(v) over v in N(u)
// below u (c4_bound_kernel): b <= kC4WaveBound one wave per root with a
// per-wave LDS hash of the endpoints; b <= kC4BlockBound one workgroup per
// root with a workgroup hash; above it the workgroup buckets the root's
// endpoints by a hash (an LDS histogram, then a scatter into an HBM scratch
// region, bucket after bucket) and counts each bucket in the LDS hash.
constexpr uint32_t kC4WaveSlots = 1024;    // per-wave hash: keys and counts, 8 KB
constexpr uint32_t kC4WaveBound = kC4WaveSlots / 2;
constexpr uint32_t kC4Slots = 8192;        // workgroup hash of the block kernel, 64 KB
constexpr uint32_t kC4BlockBound = kC4Slots / 2;
constexpr int kC4Block = 1024;
constexpr uint32_t kC4HeavySlots = 16384;  // bucket hash of the heavy kernel, 128 KB
constexpr uint32_t kC4MaxBuckets = 4096;
constexpr uint32_t kC4BucketKeys = 4096;   // mean endpoints per bucket aimed at
constexpr uint32_t kC4ScratchKeys = 12u << 20;  // per heavy workgroup: 48 MB of endpoints
constexpr uint32_t kC4BucketMul = 0x85EBCA6Bu;  // bucket hash (the tables use 2654435761)
constexpr int kC4Unroll = 4;                    // wedge stream loads in flight per lane
constexpr uint32_t kC4DirectBound = kC4HeavySlots / 2;  // heavy roots counted in one unbucketed pass
constexpr uint32_t kC4MaxProbe = 64;            // bucket table probe limit
constexpr uint32_t kAllBuckets = 0xFFFFFFFFu;

// b(u) for every vertex (one wave per vertex).
__global__ void c4_bound_kernel(const uint64_t* __restrict__ rowptr, const uint32_t* __restrict__ col,
                                const uint32_t* __restrict__ deg, uint32_t n, uint64_t* __restrict__ bound) {
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t wpb = blockDim.x >> 6;
    for (uint32_t u = blockIdx.x * wpb + (threadIdx.x >> 6); u < n; u += gridDim.x * wpb) {
        const uint64_t ub = rowptr[u];
        const uint32_t du = uint32_t(rowptr[u + 1] - ub);
        unsigned long long sum = 0;
        for (uint32_t base = 0; base < du; base += 64) {
            const uint32_t i = base + lane;
            const uint32_t v = i < du ? col[ub + i] : 0xFFFFFFFFu;
            if (v < u) sum += deg[v];
            if (__ballot(v >= u)) break;  // sorted lists: the rest ranks above u
        }
        sum = wave_sum_u64(sum);
        if (lane == 0) bound[u] = sum;
    }
}

__device__ __forceinline__ uint32_t shfl_u32(uint32_t v, uint32_t l) { return uint32_t(__shfl(int(v), int(l))); }

// The wedges u - v - w (v = N(u)[k], k < pre; w in N(v), w < u) of one root,
// streamed by one wave: groups of 64 neighbours v (k0, k0 + kstep, ...),
// whose lists are concatenated; each lane walks its own contiguous range of
// ceil(T / 64) positions of the concatenation (found once per group by a
// binary search of the list offsets), kC4Unroll loads in flight, moving to
// the next list when it reaches its offset.  The neighbours are in rank
// order, i.e. by degree, so a group's lists have similar lengths.  f(w) runs
// on the lanes holding an endpoint.  `wl` is the wave's LDS: 65 offsets
// (the last one the total) and 64 list starts (census lists hold fewer
// than 2^31 entries).
template <class F>
__device__ __forceinline__ void c4_wave_wedges(const uint64_t* __restrict__ rowptr, const uint32_t* __restrict__ col,
                                               uint64_t ub, uint32_t pre, uint32_t u, uint32_t k0, uint32_t kstep,
                                               uint32_t* wl, F&& f) {
    const uint32_t lane = threadIdx.x & 63;
    uint32_t* soff = wl;
    uint32_t* sstart = wl + 65;
    for (uint32_t k = k0; k < pre; k += kstep) {
        uint32_t vb = 0, dv = 0;
        if (k + lane < pre) {
            const uint32_t v = ldg(col, ub + k + lane);
            const uint64_t b = ldg(rowptr, v);
            vb = uint32_t(b);
            dv = uint32_t(ldg(rowptr, v + 1) - b);
        }
        uint32_t incl = dv;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const uint32_t t = uint32_t(__shfl_up(int(incl), d));
            if (lane >= uint32_t(d)) incl += t;
        }
        const uint32_t T = wave_lane(incl, 63);
        wave_lds_sync();
        soff[lane] = incl - dv;
        sstart[lane] = vb;
        if (lane == 63) soff[64] = T;
        wave_lds_sync();
        const uint32_t q = (T + 63) / 64;
        uint32_t p = min(lane * q, T);
        const uint32_t e = min(p + q, T);
        // the list holding p: the last whose offset is <= p (non-empty)
        uint32_t lo = 0;
#pragma unroll
        for (uint32_t step = 32; step; step >>= 1) {
            const uint32_t t = lo + step;
            if (t < 64 && soff[t] <= p) lo = t;
        }
        uint32_t nxt = soff[lo + 1];
        uint32_t adr = sstart[lo] - soff[lo];  // col index of position p is adr + p (mod 2^32)
        while (p < e) {
            uint32_t ws[kC4Unroll];
            const uint32_t cnt = min(uint32_t(kC4Unroll), e - p);
#pragma unroll
            for (int c = 0; c < kC4Unroll; ++c) {
                ws[c] = 0xFFFFFFFFu;
                if (uint32_t(c) < cnt) {
                    if (p == nxt) {
                        do {
                            ++lo;
                            nxt = soff[lo + 1];
                        } while (nxt == p);
                        adr = sstart[lo] - soff[lo];
                    }
                    ws[c] = ldg(col, uint64_t(adr + p));
                    ++p;
                }
            }
#pragma unroll
            for (int c = 0; c < kC4Unroll; ++c)
                if (ws[c] < u) f(ws[c]);
        }
    }
}

// Counts w into an open-addressing LDS hash of (key, count) with H slots.
__device__ __forceinline__ void c4_hash_add(uint32_t* key, uint32_t* cnt, uint32_t hmask, uint32_t hshift, uint32_t w) {
    uint32_t h = hash_slot(w, hshift);
    while (true) {
        const uint32_t k = key[h];
        if (k == w) break;
        if (k == kEmpty) {
            const uint32_t old = atomicCAS(&key[h], kEmpty, w);
            if (old == kEmpty || old == w) break;
        }
        h = (h + 1) & hmask;
    }
    atomicAdd(&cnt[h], 1u);
}

__device__ __forceinline__ void c4_geometry(uint32_t keys, uint32_t* H, uint32_t* hshift) {
    uint32_t h = 64, bits = 6;
    while (h < 2 * keys) {
        h <<= 1;
        ++bits;
    }
    *H = h;
    *hshift = 32 - bits;
}

// Roots with b(u) <= kC4WaveBound: one wave each.
__global__ __launch_bounds__(256) void census_c4_wave_kernel(const uint64_t* __restrict__ rowptr,
                                                             const uint32_t* __restrict__ col,
                                                             const uint32_t* __restrict__ tasks,
                                                             const uint64_t* __restrict__ bound, uint32_t ntasks,
                                                             uint32_t* __restrict__ cursor,
                                                             unsigned long long* __restrict__ sums) {
    __shared__ uint32_t sKey[4][kC4WaveSlots], sCnt[4][kC4WaveSlots];
    __shared__ uint32_t sList[4][129];
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t wave = threadIdx.x >> 6;
    uint32_t* key = sKey[wave];
    uint32_t* cnt = sCnt[wave];
    for (uint32_t i = lane; i < kC4WaveSlots; i += 64) {
        key[i] = kEmpty;
        cnt[i] = 0;
    }
    wave_lds_sync();
    unsigned long long acc = 0;
    TaskChunk chunk(ntasks, gridDim.x * 4, 8);
    while (true) {
        const uint32_t t = wave_next_task(chunk, cursor);
        if (t >= ntasks) break;
        const uint32_t u = wave_bcast(tasks[t]);
        const uint64_t ub = ldg(rowptr, u);
        const uint32_t du = uint32_t(ldg(rowptr, u + 1) - ub);
        const uint32_t pre = lower_bound_global(col + ub, du, u);
        uint32_t H, hshift;
        c4_geometry(uint32_t(bound[u]), &H, &hshift);
        // the table is clear: every root clears what it used while summing
        c4_wave_wedges(rowptr, col, ub, pre, u, 0, 64, sList[wave],
                       [&](uint32_t w) { c4_hash_add(key, cnt, H - 1, hshift, w); });
        wave_lds_sync();
        for (uint32_t i = lane; i < H; i += 64) {
            const uint64_t c = cnt[i];
            acc += c * (c - 1) / 2;
            key[i] = kEmpty;
            cnt[i] = 0;
        }
        wave_lds_sync();
    }
    const unsigned long long x = wave_sum_u64(acc);
    if (lane == 0 && x) atomicAdd(&sums[6], x);
}

// Roots with kC4WaveBound < b(u) <= kC4BlockBound: one workgroup each.
__global__ __launch_bounds__(kC4Block) void census_c4_block_kernel(const uint64_t* __restrict__ rowptr,
                                                                   const uint32_t* __restrict__ col,
                                                                   const uint32_t* __restrict__ tasks,
                                                                   const uint64_t* __restrict__ bound,
                                                                   uint32_t ntasks, uint32_t* __restrict__ cursor,
                                                                   unsigned long long* __restrict__ sums) {
    __shared__ uint32_t sTask;
    __shared__ uint32_t sKey[kC4Slots], sCnt[kC4Slots];
    __shared__ uint32_t sList[kC4Block / 64][129];
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t wave = threadIdx.x >> 6;
    const uint32_t nw = blockDim.x >> 6;
    unsigned long long acc = 0;
    for (uint32_t i = threadIdx.x; i < kC4Slots; i += blockDim.x) {
        sKey[i] = kEmpty;
        sCnt[i] = 0;
    }
    TaskChunk chunk(ntasks, gridDim.x, 8);
    while (true) {
        if (threadIdx.x == 0) sTask = thread_next_task(chunk, cursor);
        __syncthreads();
        const uint32_t t = wave_bcast(sTask);
        if (t >= ntasks) break;
        const uint32_t u = wave_bcast(tasks[t]);
        const uint64_t ub = ldg(rowptr, u);
        const uint32_t du = uint32_t(ldg(rowptr, u + 1) - ub);
        const uint32_t pre = lower_bound_global(col + ub, du, u);
        uint32_t H, hshift;
        c4_geometry(uint32_t(bound[u]), &H, &hshift);
        // the table is clear: every root clears what it used while summing
        c4_wave_wedges(rowptr, col, ub, pre, u, wave * 64, nw * 64, sList[wave],
                       [&](uint32_t w) { c4_hash_add(sKey, sCnt, H - 1, hshift, w); });
        __syncthreads();
        for (uint32_t i = threadIdx.x; i < H; i += blockDim.x) {
            const uint64_t c = sCnt[i];
            acc += c * (c - 1) / 2;
            sKey[i] = kEmpty;
            sCnt[i] = 0;
        }
        __syncthreads();
    }
    const unsigned long long x = wave_sum_u64(acc);
    if (lane == 0 && x) atomicAdd(&sums[6], x);
}

// Exclusive scan of a[0..P) in place by the whole workgroup; a[P] = total.
__device__ void block_exclusive_scan(uint32_t* a, uint32_t P, uint32_t* wave_tot) {
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t wave = threadIdx.x >> 6;
    const uint32_t nw = blockDim.x >> 6;
    const uint32_t per = (P + blockDim.x - 1) / blockDim.x;
    const uint32_t b0 = threadIdx.x * per;
    uint32_t local = 0;
    for (uint32_t i = 0; i < per; ++i)
        if (b0 + i < P) local += a[b0 + i];
    uint32_t incl = local;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t t = uint32_t(__shfl_up(int(incl), d));
        if (lane >= uint32_t(d)) incl += t;
    }
    if (lane == 63) wave_tot[wave] = incl;
    __syncthreads();
    uint32_t before = 0;
    for (uint32_t i = 0; i < wave; ++i) before += wave_tot[i];
    uint32_t run = before + incl - local;
    for (uint32_t i = 0; i < per; ++i)
        if (b0 + i < P) {
            const uint32_t x = a[b0 + i];
            a[b0 + i] = run;
            run += x;
        }
    if (threadIdx.x == blockDim.x - 1) {
        uint32_t tot = 0;
        for (uint32_t i = 0; i < nw; ++i) tot += wave_tot[i];
        a[P] = tot;
    }
    __syncthreads();
}

// Roots with b(u) > kC4BlockBound: one workgroup each.
//   0. b(u) <= kC4DirectBound: one pass counts every endpoint in the LDS
//      table (at most half full);
//   1. endpoints are bucketed by a hash into P ~ b(u) / kC4BucketKeys
//      buckets of a fixed capacity in the workgroup's HBM scratch, as many
//      buckets per round (one stream of the wedges) as the scratch holds,
//      and each bucket is counted in the LDS table;
//   2. if a bucket fills up, the remaining buckets take the exact path: a
//      histogram, an exclusive scan of the bucket starts, and rounds of
//      consecutive buckets whose endpoints fit the scratch.
// A bucket whose distinct endpoints overfill the table (an insert finds
// no slot within kC4MaxProbe of its home and flags it) is counted again in
// 2, 4, ... parts split by further hash bits; a bucket larger than the
// scratch is counted that way straight from the wedge stream.
__global__ __launch_bounds__(kC4Block) void census_c4_heavy_kernel(const uint64_t* __restrict__ rowptr,
                                                                   const uint32_t* __restrict__ col,
                                                                   const uint32_t* __restrict__ tasks,
                                                                   const uint64_t* __restrict__ bound,
                                                                   uint32_t ntasks, uint32_t* __restrict__ cursor,
                                                                   uint32_t* __restrict__ scratch_slab,
                                                                   uint32_t scratch_keys, uint32_t max_slots,
                                                                   uint32_t direct_bound,
                                                                   unsigned long long* __restrict__ sums) {
    __shared__ uint32_t sKey[kC4HeavySlots], sCnt[kC4HeavySlots];
    __shared__ uint32_t sOff[kC4MaxBuckets + 1];
    __shared__ uint32_t sWave[kC4Block / 64];
    __shared__ uint32_t sList[kC4Block / 64][129];
    __shared__ uint32_t sTask, sB1;
    __shared__ uint32_t sOver[2];  // overflow flag per bucket part, alternating
    uint32_t* scratch = scratch_slab + size_t(blockIdx.x) * scratch_keys;
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t wave = threadIdx.x >> 6;
    const uint32_t nw = blockDim.x >> 6;
    unsigned long long acc = 0;
    // the bucket table starts clear and every part clears what it used
    for (uint32_t i = threadIdx.x; i < kC4HeavySlots; i += blockDim.x) {
        sKey[i] = kEmpty;
        sCnt[i] = 0;
    }
    if (threadIdx.x == 0) sOver[0] = sOver[1] = 0;
    uint32_t flip = 0;
    while (true) {
        if (threadIdx.x == 0) sTask = atomicAdd(cursor, 1u);
        __syncthreads();
        const uint32_t t = wave_bcast(sTask);
        if (t >= ntasks) break;
        const uint32_t u = wave_bcast(tasks[t]);
        const uint64_t ub = ldg(rowptr, u);
        const uint32_t du = uint32_t(ldg(rowptr, u + 1) - ub);
        const uint32_t pre = lower_bound_global(col + ub, du, u);
        const uint64_t bu = bound[u];
        uint32_t P = 1, pbits = 0;
        while (P < kC4MaxBuckets && uint64_t(P) * kC4BucketKeys < bu) {
            P <<= 1;
            ++pbits;
        }
        auto bucket_of = [&](uint32_t w) { return pbits ? (w * kC4BucketMul) >> (32 - pbits) : 0u; };
        // part of w among `parts` (a power of two) inside its bucket
        auto part_of = [&](uint32_t w, uint32_t parts) {
            return parts > 1 ? ((w * kC4BucketMul) << pbits) >> (32 - __builtin_ctz(parts)) : 0u;
        };
        // Counts the endpoints of bucket b (keys[0..k) of the scratch, or the
        // wedge stream when keys is null; every endpoint for kAllBuckets) into
        // acc.  Returns false, counting nothing, when a split into more than
        // max_parts parts would be needed.
        auto count_bucket = [&](uint32_t b, const uint32_t* keys, uint32_t k, uint32_t max_parts) {
            uint32_t parts = 1;
            while (true) {
                unsigned long long bacc = 0;
                bool over = false;
                for (uint32_t part = 0; part < parts && !over; ++part) {
                    const uint32_t want = (k + parts - 1) / parts;
                    uint32_t H, hshift;
                    c4_geometry(min(want, max_slots / 2), &H, &hshift);
                    uint32_t& ovf = sOver[flip];
                    // A key is placed within kC4MaxProbe slots of its home
                    // slot or the part is flagged (the table is too full):
                    // lookups then stop there too, and a flagged part is
                    // counted again, split further.  (No occupancy counter:
                    // one LDS word taking every insert serialises them.)
                    const uint32_t max_probe = min(H, kC4MaxProbe);
                    auto add = [&](uint32_t w) {
                        uint32_t h = hash_slot(w, hshift);
                        for (uint32_t probes = 0; probes < max_probe; ++probes) {
                            const uint32_t kk = sKey[h];
                            if (kk == w) {
                                atomicAdd(&sCnt[h], 1u);
                                return;
                            }
                            if (kk == kEmpty) {
                                const uint32_t old = atomicCAS(&sKey[h], kEmpty, w);
                                if (old == kEmpty || old == w) {
                                    atomicAdd(&sCnt[h], 1u);
                                    return;
                                }
                            }
                            h = (h + 1) & (H - 1);
                        }
                        ovf = 1;
                    };
                    if (keys) {
                        // four keys per thread loaded together
                        for (uint32_t i0 = threadIdx.x; i0 < k; i0 += 4 * blockDim.x) {
                            uint32_t kw[4];
#pragma unroll
                            for (int c = 0; c < 4; ++c) {
                                const uint32_t i = i0 + c * blockDim.x;
                                kw[c] = i < k ? keys[i] : kEmpty;
                            }
#pragma unroll
                            for (int c = 0; c < 4; ++c)
                                if (kw[c] != kEmpty && part_of(kw[c], parts) == part) add(kw[c]);
                        }
                    } else {
                        c4_wave_wedges(rowptr, col, ub, pre, u, wave * 64, nw * 64, sList[wave], [&](uint32_t w) {
                            if ((b == kAllBuckets || bucket_of(w) == b) && part_of(w, parts) == part) add(w);
                        });
                    }
                    __syncthreads();
                    over = ovf != 0;
                    for (uint32_t i = threadIdx.x; i < H; i += blockDim.x) {
                        const uint64_t c = sCnt[i];
                        if (!over) bacc += c * (c - 1) / 2;
                        sKey[i] = kEmpty;
                        sCnt[i] = 0;
                    }
                    // the other flag pair was last read before the barrier
                    // that opened this part
                    if (threadIdx.x == 0) sOver[flip ^ 1] = 0;
                    flip ^= 1;
                    __syncthreads();
                }
                if (!over) {
                    acc += bacc;
                    return true;
                }
                if (parts >= max_parts) return false;
                parts *= 2;  // this split left a part above the table: split further
            }
        };
        // Roots with at most half as many wedges as the table has slots are
        // counted in one pass over the whole table (split only if a probe
        // run grows past the limit).
        if (bu <= direct_bound && count_bucket(kAllBuckets, nullptr, uint32_t(bu), 2)) continue;
        // When the scratch holds every bucket at a fixed capacity (1.5 times
        // the mean the bound allows, plus 64), one stream appends each
        // endpoint to its bucket; a bucket that fills up sends the root to
        // the exact path below, which larger roots take directly.
        const uint64_t cap64 = bu * 3 / (2 * uint64_t(P)) + 64;
        bool exact = cap64 * P > scratch_keys;
        if (!exact) {
            const uint32_t cap = uint32_t(cap64);
            for (uint32_t b = threadIdx.x; b < P; b += blockDim.x) sOff[b] = 0;
            if (threadIdx.x == 0) sB1 = 0;  // a bucket filled up
            __syncthreads();
            c4_wave_wedges(rowptr, col, ub, pre, u, wave * 64, nw * 64, sList[wave], [&](uint32_t w) {
                const uint32_t b = bucket_of(w);
                const uint32_t pos = atomicAdd(&sOff[b], 1u);
                if (pos < cap) scratch[b * cap + pos] = w;
                else sB1 = 1;
            });
            __syncthreads();
            exact = sB1 != 0;
            if (!exact)
                for (uint32_t b = 0; b < P; ++b) {
                    const uint32_t k = sOff[b];
                    if (k > 1) count_bucket(b, scratch + b * cap, k, ~0u);
                }
        }
        if (!exact) continue;
        // Exact path: histogram, exclusive scan of the bucket starts, then
        // rounds of consecutive buckets whose endpoints fit the scratch.
        __syncthreads();
        for (uint32_t b = threadIdx.x; b <= P; b += blockDim.x) sOff[b] = 0;
        __syncthreads();
        c4_wave_wedges(rowptr, col, ub, pre, u, wave * 64, nw * 64, sList[wave],
                       [&](uint32_t w) { atomicAdd(&sOff[bucket_of(w)], 1u); });
        __syncthreads();
        block_exclusive_scan(sOff, P, sWave);
        uint32_t b0 = 0;
        while (b0 < P) {
            const uint32_t base0 = sOff[b0];
            if (threadIdx.x == 0) {
                // the last b1 with sOff[b1] - base0 <= scratch (starts are sorted)
                uint32_t lo = b0, hi = P;
                while (lo < hi) {
                    const uint32_t mid = (lo + hi + 1) >> 1;
                    if (sOff[mid] - base0 <= scratch_keys) lo = mid;
                    else hi = mid - 1;
                }
                sB1 = lo;
            }
            __syncthreads();
            const uint32_t b1 = sB1;
            if (b1 == b0) {
                // one bucket above the scratch: count it from the wedge stream
                const uint32_t k = sOff[b0 + 1] - base0;
                count_bucket(b0, nullptr, k, ~0u);
                b0 += 1;
                continue;
            }
            c4_wave_wedges(rowptr, col, ub, pre, u, wave * 64, nw * 64, sList[wave], [&](uint32_t w) {
                const uint32_t b = bucket_of(w);
                if (b >= b0 && b < b1) scratch[atomicAdd(&sOff[b], 1u) - base0] = w;
            });
            __syncthreads();
            // sOff[b] now holds the end of bucket b, i.e. the start of b + 1
            for (uint32_t b = b0; b < b1; ++b) {
                const uint32_t s0 = b == b0 ? base0 : sOff[b - 1];
                const uint32_t k = sOff[b] - s0;
                if (k > 1) count_bucket(b, scratch + (s0 - base0), k, ~0u);
            }
            b0 = b1;
        }
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

// 4-cycle pass: wedge bounds, the three root lists (largest bound first),
// and the heavy kernel's endpoint scratch (one region per workgroup, one
// workgroup per CU, fewer when the regions would take more than a quarter
// of free memory).
static int prepare_c4(Ranked& R) {
    const Runtime& rt_ = rt();
    hipStream_t s = rt_.stream;
    const uint32_t n = R.n;
    KZ_HIP(hipMalloc(&R.c4_bound, 8 * std::max<size_t>(1, n)));
    if (n)
        hipLaunchKernelGGL(c4_bound_kernel, dim3(std::max(1, rt_.num_cus * 8)), dim3(256), 0, s, R.rowptr, R.col,
                           R.deg, n, R.c4_bound);
    KZ_HIP(hipGetLastError());
    std::vector<uint64_t> b(n);
    if (n) KZ_HIP(hipMemcpyAsync(b.data(), R.c4_bound, 8 * size_t(n), hipMemcpyDeviceToHost, s));
    KZ_HIP(hipStreamSynchronize(s));
    std::vector<uint32_t> wave, block, heavy;
    for (uint32_t u = 0; u < n; ++u) {
        if (b[u] < 2) continue;  // fewer than two wedges close no 4-cycle
        if (b[u] <= kC4WaveBound) wave.push_back(u);
        else if (b[u] <= kC4BlockBound) block.push_back(u);
        else heavy.push_back(u);
    }
    auto by_bound = [&](uint32_t x, uint32_t y) { return b[x] > b[y]; };
    std::stable_sort(wave.begin(), wave.end(), by_bound);
    std::stable_sort(block.begin(), block.end(), by_bound);
    std::stable_sort(heavy.begin(), heavy.end(), by_bound);
    R.n_c4_wave = uint32_t(wave.size());
    R.n_c4_block = uint32_t(block.size());
    R.n_c4_heavy = uint32_t(heavy.size());
    std::vector<uint32_t> all(wave);
    all.insert(all.end(), block.begin(), block.end());
    all.insert(all.end(), heavy.begin(), heavy.end());
    KZ_HIP(hipMalloc(&R.c4_tasks, 4 * std::max<size_t>(1, all.size())));
    if (!all.empty()) KZ_HIP(hipMemcpy(R.c4_tasks, all.data(), 4 * all.size(), hipMemcpyHostToDevice));
    R.bytes += 8 * size_t(n) + 4 * all.size();
    // KZ_C4_SCRATCH_KEYS / KZ_C4_BUCKET_SLOTS / KZ_C4_DIRECT_BOUND shrink
    // the scratch, the bucket table and the unbucketed attempt (tests use
    // them to reach every path on small graphs)
    R.c4_scratch_keys = kC4ScratchKeys;
    R.c4_max_slots = kC4HeavySlots;
    R.c4_direct_bound = kC4DirectBound;
    if (const char* e = std::getenv("KZ_C4_DIRECT_BOUND"))
        R.c4_direct_bound = uint32_t(std::min<unsigned long>(kC4DirectBound, std::strtoul(e, nullptr, 10)));
    if (const char* e = std::getenv("KZ_C4_SCRATCH_KEYS"))
        R.c4_scratch_keys = uint32_t(std::min<unsigned long>(kC4ScratchKeys, std::max<unsigned long>(1, std::strtoul(e, nullptr, 10))));
    if (const char* e = std::getenv("KZ_C4_BUCKET_SLOTS")) {
        uint32_t h = 64;
        const unsigned long want = std::strtoul(e, nullptr, 10);
        while (h < kC4HeavySlots && h < want) h <<= 1;
        R.c4_max_slots = h;
    }
    if (R.n_c4_heavy) {
        size_t free_b = 0, total_b = 0;
        KZ_HIP(hipMemGetInfo(&free_b, &total_b));
        const size_t per_wg = size_t(R.c4_scratch_keys) * 4;
        R.c4_heavy_grid = uint32_t(std::max<size_t>(
            1, std::min<size_t>(std::min<size_t>(size_t(rt_.num_cus), R.n_c4_heavy), free_b / 4 / per_wg)));
        if (hipMalloc(&R.c4_scratch, per_wg * R.c4_heavy_grid) != hipSuccess) {
            (void)hipGetLastError();
            KZ_FAIL(KZ_ERR_NOMEM, "4-cycle endpoint scratch (%zu bytes)", per_wg * R.c4_heavy_grid);
        }
        R.bytes += per_wg * R.c4_heavy_grid;
    }
    return KZ_OK;
}

// Task lists, per-edge arrays, the 4-cycle pass and the 4-clique job.
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
    KZ_HIP(hipMalloc(&R.cursor, 4 * kCursors * kCursorStride));
    KZ_HIP(hipMalloc(&R.sums, 8 * 8));
    R.bytes += 8 * (2 * std::max<uint64_t>(1, R.fnnz) + all.size() + 8);
    if (int e = prepare_c4(R)) return e;
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
    hipLaunchKernelGGL(fill_small_kernel, dim3(1), dim3(64), 0, s, R.cursor, int(kCursors * kCursorStride), R.sums, 8);
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
    if (R.n_c4_heavy)
        hipLaunchKernelGGL(census_c4_heavy_kernel, dim3(R.c4_heavy_grid), dim3(kC4Block), 0, s, R.rowptr, R.col,
                           R.c4_tasks + R.n_c4_wave + R.n_c4_block, R.c4_bound, R.n_c4_heavy,
                           R.cursor + 4 * kCursorStride, R.c4_scratch, R.c4_scratch_keys, R.c4_max_slots,
                           R.c4_direct_bound, R.sums);
    if (R.n_c4_block)
        hipLaunchKernelGGL(census_c4_block_kernel, dim3(std::min<uint32_t>(R.n_c4_block, uint32_t(rt_.num_cus) * 2)),
                           dim3(kC4Block), 0, s, R.rowptr, R.col, R.c4_tasks + R.n_c4_wave, R.c4_bound, R.n_c4_block,
                           R.cursor + 3 * kCursorStride, R.sums);
    if (R.n_c4_wave)
        hipLaunchKernelGGL(census_c4_wave_kernel,
                           dim3(std::min<uint32_t>((R.n_c4_wave + 3) / 4, uint32_t(rt_.num_cus) * 4)), dim3(256), 0, s,
                           R.rowptr, R.col, R.c4_tasks, R.c4_bound, R.n_c4_wave, R.cursor + 2 * kCursorStride, R.sums);
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
        stats->launches = 8;
        stats->tasks = uint64_t(R->nsmall) + R->nlarge + R->n_c4_wave + R->n_c4_block + R->n_c4_heavy;
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
