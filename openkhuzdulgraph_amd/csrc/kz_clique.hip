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
//     in LDS with an open-addressing hash (vertex id -> index);
//   * the induced DAG on N+(u) is built as bitmap rows: row i holds bit j
//     (j > i) iff N+(u)[j] is in N+(N+(u)[i]).  The lists N+(N+(u)[i]) of all
//     rows are read as ONE stream of aligned 16-byte groups of 4 entries: an
//     LDS prefix scan of the rows' group counts maps each lane's group index
//     to (row, group), every lane keeps several independent 16-byte loads in
//     flight and probes the hash with each entry (a wave walking one row at
//     a time is bound by dependent load latency).  Rows whose list is much
//     longer than their remaining candidates binary-search the candidates
//     in the list instead;
//   * cliques are counted with wave64 AND + popcount over the rows: k=4 sums
//     popc(row_i & row_j) over DAG edges (i, j); k=5 descends one level more;
//     k=3 counts the matches directly;
//   * tasks are bucketed by |N+(u)| with runtime bounds (autotunable): the
//     first bucket (<= 64) runs one task per wave with one u64 row per
//     candidate, the middle buckets one task per workgroup with rows in LDS,
//     and roots above the last bound ("hubs") are split over many
//     workgroups: a rows kernel builds chunks of rows into an HBM slab, then
//     a count kernel takes one row per wave;
//   * every (k, bucket) is a persistent kernel pulling tasks, largest first,
//     from a device atomic cursor; all of them are captured in one hipGraph
//     with parallel branches, replayed per run.  A batch of many graphs and
//     patterns is the same graph with tasks tagged by query.
// Per-lane uint64 partial counts are reduced per wave and added to the
// query's result with one 64-bit atomic when the wave moves to another
// query or finishes.

#include <algorithm>
#include <chrono>
#include <cstdlib>
#include <cstring>
#include <map>

#include "kz_device.h"
#include "kz_kernels.h"
#include "kz_stream.h"

namespace kz {

struct QueryRef {
    const uint64_t* rowptr;
    const uint32_t* col;
};

// A hub root split over workgroups (bucket above the last bound).
struct HubMeta {
    uint32_t u, q, d, W;
    uint64_t slab_off;  // in u64 words
};

// ------------------------------------------------------------ wave tasks --
// Roots per cursor atomic of the wave-task kernel (measured on RMAT-22:
// chunks shrinking to one root at the end of the list 39.0 ms; fixed chunks
// of 4 / 8 / 16 / 32: 37.6 / 37.0 / 36.9 / 36.8 ms).
constexpr uint32_t kWaveTaskChunk = 32;

// One task per wave, |N+(u)| <= 64.  LDS per wave: L[64], 128-slot hash,
// 64 bitmap rows (one u64 each) and the row descriptors.  Tasks are small,
// so the chain cursor -> root -> list bounds of the NEXT task is issued
// while the current one streams and counts.  Register budget: 6 waves per
// SIMD (at most 80 VGPRs; the kernel is latency-bound, and 5 waves with 96
// VGPRs measured 32.84 against 32.70 ms on RMAT-22 4-cliques, 7 32.91).
template <int K, bool QU = true>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(6))) void clique_wave_kernel(const QueryRef* __restrict__ queries,
                                                          const uint32_t* __restrict__ task_v,
                                                          const uint32_t* __restrict__ task_q, uint32_t ntasks,
                                                          uint32_t* __restrict__ cursor,
                                                          unsigned long long* __restrict__ results) {
    constexpr int WPB = 4;
    constexpr uint32_t H = 128;
    __shared__ uint32_t sL[WPB][64];
    __shared__ __align__(8) uint32_t sKey[WPB][H];
    __shared__ __align__(4) uint16_t sVal[WPB][H];
    __shared__ uint32_t sFb[WPB][32];  // 1024-bit filter: 16 bits per id
    __shared__ unsigned long long sRow[WPB][64];
    __shared__ unsigned long long sRb[WPB][64];
    __shared__ uint32_t sRp[WPB][64];
    __shared__ __align__(8) uint32_t sQ[QU ? WPB : 1][128];  // probe queues
    const uint32_t lane = threadIdx.x & 63;
    const int w = threadIdx.x >> 6;
    uint32_t* L = sL[w];
    uint32_t* hk = sKey[w];
    uint16_t* hv = sVal[w];
    const RootSet rs{hk, hv, H - 1, 25, sFb[w], 22};
    unsigned long long* rows = sRow[w];
    unsigned long long* rb = sRb[w];
    uint32_t* rp = sRp[w];
    unsigned long long acc = 0;
    uint32_t qcur = 0xFFFFFFFFu;
    const uint32_t workers = gridDim.x * WPB;
    TaskChunk chunk(ntasks, workers, 2, kWaveTaskChunk, kWaveTaskChunk);
    auto fetch = [&]() { return wave_next_task(chunk, cursor); };
    // current task: index, query, list start and length, graph arrays
    uint32_t t = fetch(), q = 0, d = 0;
    uint64_t ub = 0;
    const uint64_t* rowptr = nullptr;
    const uint32_t* col = nullptr;
    if (t < ntasks) {
        q = task_q ? task_q[t] : 0u;
        rowptr = queries[q].rowptr;
        col = queries[q].col;
        const uint32_t u = task_v[t];
        ub = ldg(rowptr, u);
        d = uint32_t(ldg(rowptr, u + 1) - ub);
    }
    while (t < ntasks) {
        if (q != qcur) {
            if (qcur != 0xFFFFFFFFu) flush_query(acc, qcur, results);
            qcur = q;
        }
        const uint32_t x = lane < d ? ldg(col, ub + lane) : 0xFFFFFFFFu;
        const uint32_t tn = fetch();
        L[lane] = x;
        hk[lane] = kEmpty;
        hk[lane + 64] = kEmpty;
        if (lane < 32) rs.fb[lane] = 0;
        rows[lane] = 0;
        // row descriptors: lane r owns row r
        // (rows >= d - 1 are empty, so rp[d - 1] is the group total)
        uint32_t groups = 0, len = 0;
        uint64_t desc = 0;
        bool skewed = false;
        if (lane + 1 < d) skewed = row_desc(rowptr, x, d - 1 - lane, &desc, &groups, &len);
        rb[lane] = desc;
        uint32_t excl;
        const uint32_t G = wave_exclusive_scan(groups, &excl);
        rp[lane] = excl;
        wave_lds_sync();
        if (lane < d) root_insert(rs, x, uint16_t(lane));
        // next task's query and root (a cursor atomic has returned by now)
        uint32_t qn = q, un = 0;
        if (tn < ntasks) {
            qn = task_q ? task_q[tn] : 0u;
            un = task_v[tn];
        }
        wave_lds_sync();
        stream_groups<kGroups, (K >= 4), QU>(
            col, rb, rp, d - 1, G, lane, 64, rs,
            [&](uint32_t r, uint32_t j, uint64_t) {
                if (K == 3) acc += 1;
                else atomicOr(&rows[r], 1ull << j);
            },
            nullptr, QU ? sQ[QU ? w : 0] : nullptr);
        // next task's list bounds
        const uint64_t* rowptr_n = rowptr;
        const uint32_t* col_n = col;
        uint64_t ubn = 0;
        uint32_t dn = 0;
        if (tn < ntasks) {
            if (qn != q) {
                rowptr_n = queries[qn].rowptr;
                col_n = queries[qn].col;
            }
            ubn = ldg(rowptr_n, un);
            dn = uint32_t(ldg(rowptr_n, un + 1) - ubn);
        }
        // skewed rows: their (row, candidate) pairs spread over the lanes
        if (__ballot(skewed))
            search_skewed(col, rb, L, d, 0, d - 1, 0, 1, [&](uint32_t r, uint32_t j, uint64_t) {
                if (K == 3) acc += 1;
                else atomicOr(&rows[r], 1ull << j);
            });
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
        t = tn;
        q = qn;
        rowptr = rowptr_n;
        col = col_n;
        ub = ubn;
        d = dn;
    }
    if (qcur != 0xFFFFFFFFu) flush_query(acc, qcur, results);
}

// Word kw of row i of a word-major bitmap is rows[word_base(kw) + i]: kw *
// ws, or, packed (TRI), 32 kw (kw + 1).  A root's DAG rows are strictly
// upper triangular (row i has bits j > i only), so word kw of rows i >=
// 64 (kw + 1) - 1 is always zero: the packed form keeps rows i < 64 (kw +
// 1) of word kw (all d rows of the last), tri_words(d) words in all.
template <bool TRI>
__device__ __forceinline__ uint32_t word_base(uint32_t kw, uint32_t ws) {
    return TRI ? (kw * (kw + 1)) << 5 : kw * ws;
}
__host__ __device__ inline uint32_t tri_words(uint32_t d) {
    const uint32_t W = (d + 63) >> 6;
    return W ? 32 * (W - 1) * W + d : 0u;
}

// 5-cliques through DAG rows i < j (bit j of row i set), word kj = (j+1)/64
// on: with C = row_i & row_j (ids > j), the adjacent pairs l < m of C, i.e.
// the sum over l in C of |C & row_l|.  C's wd <= WMAX words stay in
// registers, so each l costs one row load per word instead of three.  TRI:
// packed rows (word_base), whose word offsets are then constants.
template <uint32_t WMAX, bool TRI = false>
__device__ __forceinline__ uint32_t count5_pairs(const unsigned long long* rows, uint32_t ws, uint32_t wd,
                                                 const unsigned long long* ri, const unsigned long long* rj,
                                                 uint32_t kj) {
    unsigned long long c[WMAX];
#pragma unroll
    for (uint32_t k = 0; k < WMAX; ++k)
        c[k] = (k >= kj && k < wd) ? ri[word_base<TRI>(k, ws)] & rj[word_base<TRI>(k, ws)] : 0ull;
    uint32_t n = 0;  // at most |C|^2 / 2 < 2^19 for wd <= 16
#pragma unroll
    for (uint32_t kk = 0; kk < WMAX; ++kk) {
        unsigned long long m = c[kk];
        while (m) {
            const unsigned long long* rl = rows + (kk * 64 + uint32_t(__builtin_ctzll(m)));
            m &= m - 1;
            // row l holds ids > l only: its words below kk are empty
#pragma unroll
            for (uint32_t k3 = kk; k3 < WMAX; ++k3)
                if (k3 < wd) n += uint32_t(__popcll(c[k3] & rl[word_base<TRI>(k3, ws)]));
        }
    }
    return n;
}

// Counts the k-cliques of one root's DAG rows i = first, first + stride, ...
// (wave-cooperative: lanes take the set bits j of row i).  The bitmap is
// word-major: word kk of row i is rows[kk * ws + i], so the lanes' rows j
// sit in consecutive words (no LDS bank conflicts) and row i is a broadcast.
// With ctr (an LDS counter zeroed before a barrier) rows are taken one
// at a time from it instead (first and stride unused): rows come largest
// first, so the waves finish together.  kRegC: 5-clique rows of at most
// 16 words keep row_i & row_j in registers (count5_pairs).
template <int K, bool kRegC = true, bool TRI = false>
__device__ __forceinline__ void count_rows(const unsigned long long* rows, uint32_t ws, uint32_t d, uint32_t first,
                                           uint32_t stride, unsigned long long& acc, uint32_t* ctr = nullptr) {
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t wd = (d + 63) >> 6;
    auto next = [&](uint32_t i) -> uint32_t {
        if (!ctr) return i + stride;
        uint32_t t = 0;
        if (lane == 0) t = atomicAdd(ctr, 1u);
        return wave_bcast(t);
    };
    for (uint32_t i = ctr ? next(0) : first; i + 2 < d; i = next(i)) {
        const unsigned long long* ri = rows + i;
        const uint32_t k_lo = (i + 1) >> 6;
        uint32_t total = 0;
        for (uint32_t k = k_lo; k < wd; ++k) total += __popcll(ri[word_base<TRI>(k, ws)]);
        total = wave_bcast(total);  // row i is the same in every lane
        for (uint32_t base = 0; base < total; base += 64) {
            const uint32_t r = base + lane;
            if (r >= total) break;
            uint32_t k = k_lo, rem = r;
            unsigned long long wv = ri[word_base<TRI>(k, ws)];
            for (uint32_t c = __popcll(wv); rem >= c; c = __popcll(wv)) {
                rem -= c;
                wv = ri[word_base<TRI>(++k, ws)];
            }
            const uint32_t j = k * 64 + select_bit(wv, rem);
            const unsigned long long* rj = rows + j;
            const uint32_t kj = (j + 1) >> 6;
            if (K == 4) {
                for (uint32_t kk = kj; kk < wd; ++kk) acc += __popcll(ri[word_base<TRI>(kk, ws)] & rj[word_base<TRI>(kk, ws)]);
            } else if (kRegC && wd <= 16) {  // d <= 1024: C in registers
                if (wd <= 2) acc += count5_pairs<2, TRI>(rows, ws, wd, ri, rj, kj);
                else if (wd <= 8) acc += count5_pairs<8, TRI>(rows, ws, wd, ri, rj, kj);
                else acc += count5_pairs<16, TRI>(rows, ws, wd, ri, rj, kj);
            } else {
                for (uint32_t kk = kj; kk < wd; ++kk) {
                    unsigned long long m = ri[word_base<TRI>(kk, ws)] & rj[word_base<TRI>(kk, ws)];
                    while (m) {
                        const uint32_t l = kk * 64 + __builtin_ctzll(m);
                        m &= m - 1;
                        const unsigned long long* rl = rows + l;
                        for (uint32_t k3 = (l + 1) >> 6; k3 < wd; ++k3)
                            acc += __popcll(ri[word_base<TRI>(k3, ws)] & rj[word_base<TRI>(k3, ws)] & rl[word_base<TRI>(k3, ws)]);
                    }
                }
            }
        }
    }
}

// 5-cliques of one root's DAG rows in LDS (d <= 1024 rows, word-major with
// stride d, or packed: TRI), the workgroup's lanes taking its edges (i, j)
// 64 at a time from ctr (an LDS counter zeroed before a barrier) in row
// order, so that rows with fewer than 64 set bits leave no lane idle.  pre:
// d + 1 u32 of free LDS (the row prefix of edges); tmp: one u32 per wave.
// Block-wide.
template <bool TRI = false>
__device__ __forceinline__ void count5_edges(const unsigned long long* rows, uint32_t d, uint32_t* pre,
                                             uint32_t* tmp, uint32_t* ctr, unsigned long long& acc) {
    const uint32_t wd = (d + 63) >> 6;
    for (uint32_t i = threadIdx.x; i <= d; i += blockDim.x) {
        uint32_t c = 0;
        if (i < d)
            for (uint32_t k = (i + 1) >> 6; k < wd; ++k) c += uint32_t(__popcll(rows[word_base<TRI>(k, d) + i]));
        pre[i] = c;
    }
    __syncthreads();
    const uint32_t E = block_exclusive_scan(pre, d + 1, tmp);
    const uint32_t lane = threadIdx.x & 63;
    uint32_t base = 0;
    if (lane == 0) base = atomicAdd(ctr, 64u);
    base = wave_bcast(base);
    while (base < E) {
        uint32_t nb = 0;
        if (lane == 0) nb = atomicAdd(ctr, 64u);
        const uint32_t e = base + lane;
        if (e < E) {
            const uint32_t i = row_in(pre, 0, d, e);  // the row holding edge e
            const unsigned long long* ri = rows + i;
            uint32_t rem = e - pre[i], k = (i + 1) >> 6;
            unsigned long long wv = ri[word_base<TRI>(k, d)];
            for (uint32_t c = uint32_t(__popcll(wv)); rem >= c; c = uint32_t(__popcll(wv))) {
                rem -= c;
                wv = ri[word_base<TRI>(++k, d)];
            }
            const uint32_t j = k * 64 + select_bit(wv, rem);
            const uint32_t kj = (j + 1) >> 6;
            if (wd <= 2) acc += count5_pairs<2, TRI>(rows, d, wd, ri, rows + j, kj);
            else if (wd <= 8) acc += count5_pairs<8, TRI>(rows, d, wd, ri, rows + j, kj);
            else acc += count5_pairs<16, TRI>(rows, d, wd, ri, rows + j, kj);
        }
        base = wave_bcast(nb);
    }
}

// ------------------------------------------------------- MFMA 4-cliques --
// With A the d x d 0/1 matrix of the root's DAG (A[i][j] = bit j of row i,
// j > i), the 4-cliques at the root are sum_{i<j} A[i][j] * |row_i & row_j|
// = sum_{i,j} A[i][j] * (A A^T)[i][j]: an int8 GEMM masked by A.  The
// workgroup's waves take 32 x 32 tiles (I, J), J >= I, of A A^T, one
// v_mfma_i32_32x32x32_i8 per 32 columns; A A^T[i][j] only gathers columns
// l > j >= 32 J, so column blocks start at J.  Operands come straight from
// the LDS bitmap: lane l holds row (or column) l % 32 of the tile and the 16
// bits (l / 32) of the 32-column block, expanded to 16 bytes of 0/1; the
// k order inside a block is the same for both operands, so any order works.
// Accumulator r of lane l holds tile row (r / 4) * 8 + (l / 32) * 4 + r % 4,
// column l % 32.
typedef int kz_v4i __attribute__((ext_vector_type(4)));
typedef int kz_v16i __attribute__((ext_vector_type(16)));

// Bits off .. off + 15 of w as 16 bytes of 0 or 1 (bit off + t in byte t):
// each nibble is a bit-field extract, a 24-bit multiply spreads its bits 7
// apart and the mask keeps one per byte.
__device__ __forceinline__ kz_v4i bits_to_bytes(uint32_t w, uint32_t off) {
    kz_v4i r;
#pragma unroll
    for (int q = 0; q < 4; ++q)
        r[q] = int(__umul24(__builtin_amdgcn_ubfe(w, off + 4u * uint32_t(q), 4u), 0x00204081u) & 0x01010101u);
    return r;
}

// This lane's share of sum_r A[i_r][j] * c[r] over its 16 rows i_r of tile
// row block I (accumulator r of lane l holds row (r / 4) * 8 + (l / 32) * 4 +
// r % 4) and its column j: A[i_r][j] read as the 32-bit half of each row's
// word that holds bit j (entries of A A^T are at most d < 2^24: 24-bit
// multiply-adds).
template <bool TRI>
__device__ __forceinline__ uint32_t mask_tile(const unsigned long long* rows, uint32_t ws, uint32_t d, uint32_t I,
                                              uint32_t j, const kz_v16i& c) {
    const uint32_t lane = threadIdx.x & 63;
    uint32_t s = 0;
    if (j < d) {
        const uint32_t* cw = reinterpret_cast<const uint32_t*>(rows + word_base<TRI>(j >> 6, ws)) + ((j >> 5) & 1u);
        const uint32_t b5 = j & 31u;
        const uint32_t ib = 32 * I + (lane >> 5) * 4;
        if (32 * I + 32 <= d) {  // every row of the tile exists (wave-uniform)
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const uint32_t ir = ib + (r >> 2) * 8 + (r & 3);
                s = __umul24(__builtin_amdgcn_ubfe(cw[2 * ir], b5, 1u), uint32_t(c[r])) + s;
            }
        } else {
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const uint32_t ir = ib + (r >> 2) * 8 + (r & 3);
                const uint32_t a = ir < d ? __builtin_amdgcn_ubfe(cw[2 * ir], b5, 1u) : 0u;
                s = __umul24(a, uint32_t(c[r])) + s;
            }
        }
    }
    return s;
}

// One 32 x 32 tile (I, J >= I) of A A^T masked by A: this lane's share of
// sum A[i][j] * (A A^T)[i][j] over the tile.  rows is a word-major bitmap
// (word kk of row i at rows[word_base<TRI>(kk, ws) + i]) in LDS or in HBM.
// Every row a tile reads exists in the packed form: rows below 32 (J + 1)
// from words kk >= J / 2, and rows clamped to d - 1 only in the last word.
// Call with every lane of the wave.
template <bool TRI = false>
__device__ __forceinline__ uint32_t mfma4_tile(const unsigned long long* rows, uint32_t ws, uint32_t d, uint32_t I,
                                               uint32_t J) {
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t nb = (d + 31) >> 5;
    const uint32_t i = 32 * I + (lane & 31), j = 32 * J + (lane & 31);
    // rows past d read row d - 1: their products land in tile rows and
    // columns that the mask below leaves out
    const uint32_t ic = i < d ? i : d - 1, jc = j < d ? j : d - 1;
    const uint32_t hs = 16 * (lane >> 5);
    // one 64-bit word = two 32-column blocks, read once, the next word's
    // reads issued before this one's products; word J / 2 may start at its
    // high block (J odd), the last word may end after its low block (nb odd)
    kz_v16i c;
#pragma unroll
    for (int r = 0; r < 16; ++r) c[r] = 0;
    const uint32_t wend = (nb + 1) >> 1;
    uint32_t kw = J >> 1;
    unsigned long long wi = rows[word_base<TRI>(kw, ws) + ic], wj = rows[word_base<TRI>(kw, ws) + jc];
    if (J & 1u) {
        const uint32_t hi_i = uint32_t(wi >> 32), hi_j = uint32_t(wj >> 32);
        ++kw;
        if (kw < wend) {
            wi = rows[word_base<TRI>(kw, ws) + ic];
            wj = rows[word_base<TRI>(kw, ws) + jc];
        }
        c = __builtin_amdgcn_mfma_i32_32x32x32_i8(bits_to_bytes(hi_i, hs), bits_to_bytes(hi_j, hs), c, 0, 0, 0);
    }
    while (2 * kw + 1 < nb) {
        const uint32_t lo_i = uint32_t(wi), hi_i = uint32_t(wi >> 32);
        const uint32_t lo_j = uint32_t(wj), hi_j = uint32_t(wj >> 32);
        ++kw;
        if (kw < wend) {
            wi = rows[word_base<TRI>(kw, ws) + ic];
            wj = rows[word_base<TRI>(kw, ws) + jc];
        }
        c = __builtin_amdgcn_mfma_i32_32x32x32_i8(bits_to_bytes(lo_i, hs), bits_to_bytes(lo_j, hs), c, 0, 0, 0);
        c = __builtin_amdgcn_mfma_i32_32x32x32_i8(bits_to_bytes(hi_i, hs), bits_to_bytes(hi_j, hs), c, 0, 0, 0);
    }
    if (2 * kw < nb)
        c = __builtin_amdgcn_mfma_i32_32x32x32_i8(bits_to_bytes(uint32_t(wi), hs), bits_to_bytes(uint32_t(wj), hs), c,
                                                  0, 0, 0);
    return mask_tile<TRI>(rows, ws, d, I, j, c);
}

// Four 32 x 32 tiles (2 bI + a, 2 bJ + b), a, b in {0, 1}, of A A^T masked
// by A in one pass over the 64-bit words kw >= bJ (block (bI, bJ), bJ >=
// bI, of the W x W word blocks).  Each word's operands, the rows of the two
// I tiles and of the two J tiles, are expanded to bytes once and feed four
// MFMAs per 32-column half: 12 VALU of expansion per MFMA instead of 24, and
// half the row reads.  Tile 2 bJ + 1 also takes the low half of word bJ,
// whose products are zero (row j has no bit l <= j).  Tiles below the
// diagonal (2 bI + 1 > 2 bJ) or past the last row are computed on clamped
// rows and left out; a clamped row, d - 1, is empty.  The accumulators take
// 64 registers: for kernels with a budget of at least 128.
template <bool TRI>
__device__ __forceinline__ uint32_t mfma4_block(const unsigned long long* rows, uint32_t ws, uint32_t d, uint32_t bI,
                                                uint32_t bJ) {
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t nb = (d + 31) >> 5, W = (d + 63) >> 6;
    const uint32_t i0 = 64 * bI + (lane & 31), j0 = 64 * bJ + (lane & 31);
    const uint32_t ic0 = i0 < d ? i0 : d - 1, ic1 = i0 + 32 < d ? i0 + 32 : d - 1;
    const uint32_t jc0 = j0 < d ? j0 : d - 1, jc1 = j0 + 32 < d ? j0 + 32 : d - 1;
    const uint32_t hs = 16 * (lane >> 5);
    kz_v16i c00, c01, c10, c11;
#pragma unroll
    for (int r = 0; r < 16; ++r) c00[r] = c01[r] = c10[r] = c11[r] = 0;
    uint32_t kw = bJ;
    unsigned long long a0 = rows[word_base<TRI>(kw, ws) + ic0], a1 = rows[word_base<TRI>(kw, ws) + ic1];
    unsigned long long b0 = rows[word_base<TRI>(kw, ws) + jc0], b1 = rows[word_base<TRI>(kw, ws) + jc1];
    while (kw < W) {
        const uint32_t kn = kw + 1;
        unsigned long long na0 = 0, na1 = 0, nb0 = 0, nb1 = 0;
        if (kn < W) {
            const uint32_t base = word_base<TRI>(kn, ws);
            na0 = rows[base + ic0];
            na1 = rows[base + ic1];
            nb0 = rows[base + jc0];
            nb1 = rows[base + jc1];
        }
        {
            const kz_v4i A0 = bits_to_bytes(uint32_t(a0), hs), A1 = bits_to_bytes(uint32_t(a1), hs);
            const kz_v4i B0 = bits_to_bytes(uint32_t(b0), hs), B1 = bits_to_bytes(uint32_t(b1), hs);
            c00 = __builtin_amdgcn_mfma_i32_32x32x32_i8(A0, B0, c00, 0, 0, 0);
            c01 = __builtin_amdgcn_mfma_i32_32x32x32_i8(A0, B1, c01, 0, 0, 0);
            c10 = __builtin_amdgcn_mfma_i32_32x32x32_i8(A1, B0, c10, 0, 0, 0);
            c11 = __builtin_amdgcn_mfma_i32_32x32x32_i8(A1, B1, c11, 0, 0, 0);
        }
        if (2 * kw + 1 < nb) {  // the word's high half holds columns below d
            const kz_v4i A0 = bits_to_bytes(uint32_t(a0 >> 32), hs), A1 = bits_to_bytes(uint32_t(a1 >> 32), hs);
            const kz_v4i B0 = bits_to_bytes(uint32_t(b0 >> 32), hs), B1 = bits_to_bytes(uint32_t(b1 >> 32), hs);
            c00 = __builtin_amdgcn_mfma_i32_32x32x32_i8(A0, B0, c00, 0, 0, 0);
            c01 = __builtin_amdgcn_mfma_i32_32x32x32_i8(A0, B1, c01, 0, 0, 0);
            c10 = __builtin_amdgcn_mfma_i32_32x32x32_i8(A1, B0, c10, 0, 0, 0);
            c11 = __builtin_amdgcn_mfma_i32_32x32x32_i8(A1, B1, c11, 0, 0, 0);
        }
        a0 = na0;
        a1 = na1;
        b0 = nb0;
        b1 = nb1;
        kw = kn;
    }
    const uint32_t I0 = 2 * bI, J0 = 2 * bJ;
    uint32_t s = mask_tile<TRI>(rows, ws, d, I0, j0, c00);
    if (J0 + 1 < nb) s += mask_tile<TRI>(rows, ws, d, I0, j0 + 32, c01);
    if (I0 + 1 < nb && I0 + 1 <= J0) s += mask_tile<TRI>(rows, ws, d, I0 + 1, j0, c10);
    if (J0 + 1 < nb && I0 + 1 < nb) s += mask_tile<TRI>(rows, ws, d, I0 + 1, j0 + 32, c11);
    return s;
}

// (I, J) of tile t = J (J + 1) / 2 + I, I <= J.
__device__ __forceinline__ void tile_coords(uint32_t t, uint32_t* I, uint32_t* J) {
    uint32_t j = uint32_t((sqrtf(8.0f * float(t) + 1.0f) - 1.0f) * 0.5f);
    if (j * (j + 1) / 2 > t) --j;
    if ((j + 1) * (j + 2) / 2 <= t) ++j;
    *J = j;
    *I = t - j * (j + 1) / 2;
}

// Tiles go to the workgroup's waves from a counter in LDS (ctr, zeroed
// before a barrier), the next tile's atomic issued ahead: tile t = J (J + 1)
// / 2 + I, so the costly tiles (small J: the most 64-bit words) come first
// and the waves finish together.
// Two 32 x 32 tiles (I, 2 bJ) and (I, 2 bJ + 1), I <= 2 bJ + 1, of A A^T
// masked by A in one pass over the words kw >= bJ: the row block I operand
// is expanded once per half for two MFMAs (18 VALU of expansion per MFMA
// instead of 24); two accumulators (32 registers).  Zero products and
// clamped rows as in mfma4_block.
template <bool TRI>
__device__ __forceinline__ uint32_t mfma4_pair(const unsigned long long* rows, uint32_t ws, uint32_t d, uint32_t I,
                                               uint32_t bJ) {
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t nb = (d + 31) >> 5, W = (d + 63) >> 6;
    const uint32_t i = 32 * I + (lane & 31), j0 = 64 * bJ + (lane & 31);
    const uint32_t ic = i < d ? i : d - 1;
    const uint32_t jc0 = j0 < d ? j0 : d - 1, jc1 = j0 + 32 < d ? j0 + 32 : d - 1;
    const uint32_t hs = 16 * (lane >> 5);
    kz_v16i c0, c1;
#pragma unroll
    for (int r = 0; r < 16; ++r) c0[r] = c1[r] = 0;
    uint32_t kw = bJ;
    unsigned long long a = rows[word_base<TRI>(kw, ws) + ic];
    unsigned long long b0 = rows[word_base<TRI>(kw, ws) + jc0], b1 = rows[word_base<TRI>(kw, ws) + jc1];
    while (kw < W) {
        const uint32_t kn = kw + 1;
        unsigned long long na = 0, nb0 = 0, nb1 = 0;
        if (kn < W) {
            const uint32_t base = word_base<TRI>(kn, ws);
            na = rows[base + ic];
            nb0 = rows[base + jc0];
            nb1 = rows[base + jc1];
        }
        {
            const kz_v4i A = bits_to_bytes(uint32_t(a), hs);
            c0 = __builtin_amdgcn_mfma_i32_32x32x32_i8(A, bits_to_bytes(uint32_t(b0), hs), c0, 0, 0, 0);
            c1 = __builtin_amdgcn_mfma_i32_32x32x32_i8(A, bits_to_bytes(uint32_t(b1), hs), c1, 0, 0, 0);
        }
        if (2 * kw + 1 < nb) {
            const kz_v4i A = bits_to_bytes(uint32_t(a >> 32), hs);
            c0 = __builtin_amdgcn_mfma_i32_32x32x32_i8(A, bits_to_bytes(uint32_t(b0 >> 32), hs), c0, 0, 0, 0);
            c1 = __builtin_amdgcn_mfma_i32_32x32x32_i8(A, bits_to_bytes(uint32_t(b1 >> 32), hs), c1, 0, 0, 0);
        }
        a = na;
        b0 = nb0;
        b1 = nb1;
        kw = kn;
    }
    uint32_t s = 0;
    if (I <= 2 * bJ) s = mask_tile<TRI>(rows, ws, d, I, j0, c0);
    if (2 * bJ + 1 < nb) s += mask_tile<TRI>(rows, ws, d, I, j0 + 32, c1);
    return s;
}

// (I, bJ) of pair t: pairs come by bJ, I = 0 .. 2 bJ + 1 (0 .. nb - 1 for
// the last word), so pair t of bJ < W - 1 is t = bJ (bJ + 1) + I.
__device__ __forceinline__ void pair_coords(uint32_t t, uint32_t W, uint32_t* I, uint32_t* bJ) {
    if (t >= (W - 1) * W) {
        *bJ = W - 1;
        *I = t - (W - 1) * W;
        return;
    }
    uint32_t b = uint32_t((sqrtf(4.0f * float(t) + 1.0f) - 1.0f) * 0.5f);
    if (b * (b + 1) > t) --b;
    if ((b + 1) * (b + 2) <= t) ++b;
    *bJ = b;
    *I = t - b * (b + 1);
}

// BLK: tile pairs (mfma4_pair) instead of single tiles, for kernels whose
// register budget holds their two accumulators.
template <bool TRI, bool BLK = false>
__device__ __forceinline__ void count_rows_mfma4(const unsigned long long* rows, uint32_t ws, uint32_t d,
                                                 uint32_t* ctr, unsigned long long& acc) {
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t nb = (d + 31) >> 5, W = (d + 63) >> 6;
    const uint32_t ntiles = BLK ? (W - 1) * W + nb : nb * (nb + 1) / 2;
    uint32_t tile = 0;
    if (lane == 0) tile = atomicAdd(ctr, 1u);
    tile = wave_bcast(tile);
    while (tile < ntiles) {
        uint32_t tn = 0;
        if (lane == 0) tn = atomicAdd(ctr, 1u);
        uint32_t I, J;
        if (BLK) {
            pair_coords(tile, W, &I, &J);
            acc += mfma4_pair<TRI>(rows, ws, d, I, J);
        } else {
            tile_coords(tile, &I, &J);
            acc += mfma4_tile<TRI>(rows, ws, d, I, J);
        }
        tile = wave_bcast(tn);
    }
}

// Sets bit j > r of row r in a word-major LDS bitmap (word kk of row r at
// rows[word_base<TRI>(kk, ws) + r], ws < 2^24).  The word index is one
// v_mad_u32_u24 (written out: the compiler otherwise distributes the byte
// scaling over the multiply and the add).
template <bool TRI = false>
__device__ __forceinline__ void set_row_bit(unsigned long long* rows, uint32_t ws, uint32_t r, uint32_t j) {
    uint32_t w;
    const uint32_t kw = j >> 6;
    if (TRI) asm("v_mad_u32_u24 %0, %1, %2, %3" : "=v"(w) : "v"(kw), "v"((kw << 5) + 32u), "v"(r));
    else asm("v_mad_u32_u24 %0, %1, %2, %3" : "=v"(w) : "v"(kw), "s"(ws), "v"(r));
    atomicOr(rows + w, 1ull << (j & 63));
}

// Builds rows [r0, r1) of the DAG of root with list L (d ids, set rs) into
// rows_out (word kk of row r at rows_out[kk * ws + r - r0], zeroed by the
// caller), streaming the neighbour lists of all rows as 16-byte groups.  rb holds
// (r1 - r0) row descriptors and rp (r1 - r0 + 1) group prefixes, both LDS
// scratch; tmp holds one u32 per wave and ctr one u32 (the stream's span
// counter).  Block-wide; NG groups of 4 entries per lane and span; QU:
// queued probes, q the workgroup's queues (64 uint2 per wave); TRI: packed
// rows (word_base).
template <int K, int NG = kGroups, bool QU = false, bool TRI = false>
__device__ __forceinline__ void build_rows_block(const uint64_t* __restrict__ rowptr, const uint32_t* __restrict__ col,
                                                 const uint32_t* L, uint32_t d, const RootSet& rs, uint32_t r0, uint32_t r1,
                                                 unsigned long long* rows_out, uint32_t ws, unsigned long long* rb,
                                                 uint32_t* rp, uint32_t* tmp, uint32_t* ctr,
                                                 unsigned long long& acc, uint32_t* q = nullptr) {
    const uint32_t nr = r1 - r0;
    if (threadIdx.x == 0) *ctr = 0;
    for (uint32_t r = threadIdx.x; r <= nr; r += blockDim.x) {
        const uint32_t i = r0 + r;
        uint32_t groups = 0, len = 0;
        uint64_t desc = 0;
        if (r < nr && i + 1 < d) row_desc(rowptr, L[i], d - 1 - i, &desc, &groups, &len);
        if (r < nr) rb[r] = desc;
        rp[r] = groups;
    }
    __syncthreads();
    const uint32_t G = block_exclusive_scan(rp, nr + 1, tmp);
    stream_groups<NG, (K >= 4), QU>(
        col, rb, rp, nr, G, threadIdx.x, blockDim.x, rs,
        [&](uint32_t r, uint32_t j, uint64_t) {
            if (K == 3) acc += 1;
            else set_row_bit<TRI>(rows_out, ws, r, j);
        },
        ctr, QU ? q + (threadIdx.x >> 6) * 128 : nullptr);
    // skewed rows: each wave searches the candidates of its own rows
    search_skewed(col, rb, L, d, r0, nr, threadIdx.x >> 6, blockDim.x >> 6, [&](uint32_t r, uint32_t j, uint64_t) {
        if (K == 3) acc += 1;
        else set_row_bit<TRI>(rows_out, ws, r, j);
    });
    __syncthreads();
}

// ----------------------------------------------------------- block tasks --
// One task per workgroup, rows in LDS.  Dynamic LDS: [rows Dmax*W u64 (k >= 4;
// TRI, 4-cliques only: packed, tri_words(Dmax) u64)]
// [row descriptors Dmax u64][probe queues 64 uint2 per wave (QU)][hash keys
// H u32][L Dmax u32][group prefix Dmax+1 u32][filter FW u32][hash vals H u16].
// QU (queued probes, where the LDS leaves room for them without costing a
// workgroup per CU): the stream's filter passes are probed 64 at a time.  Register budget: WPE
// waves per SIMD.  The default, 6 (at most 80 VGPRs), is the occupancy the
// LDS of the middle buckets allows; there the latency-bound count phase
// loses more to fewer waves than the stream gains from more registers.
// Buckets whose LDS leaves at most 4 waves per SIMD (4-cliques on the
// 512-1024 bucket) use WPE = 4 and stream NG = 3 groups per lane and span
// with the registers that frees (RMAT-22 4-clique count 34.97 -> 34.76 ms);
// those whose LDS leaves 7 or more (the 64-128 bucket, latency-bound) use
// WPE = 7 (72 VGPRs; 32.70 -> 32.43 ms, 8 waves no better), and so do
// 3-cliques, which keep no bitmap (25.81 -> 25.46 ms).
template <int K, int NG = kGroups, int WPE = 6, bool QU = false, bool TRI = false>
__global__ __launch_bounds__(1024) __attribute__((amdgpu_waves_per_eu(WPE))) void clique_block_kernel(const QueryRef* __restrict__ queries,
                                                           const uint32_t* __restrict__ task_v,
                                                           const uint32_t* __restrict__ task_q, uint32_t ntasks,
                                                           uint32_t* __restrict__ cursor,
                                                           unsigned long long* __restrict__ results, uint32_t Dmax,
                                                           uint32_t W, uint32_t H, uint32_t hshift, uint32_t FW,
                                                           uint32_t fshift) {
    extern __shared__ __align__(16) unsigned char smem[];
    unsigned char* p = smem;
    unsigned long long* rows = nullptr;
    if (K >= 4) {
        rows = reinterpret_cast<unsigned long long*>(p);
        p += size_t(TRI ? tri_words(Dmax) : Dmax * W) * 8;
    }
    unsigned long long* rb = reinterpret_cast<unsigned long long*>(p);
    p += size_t(Dmax) * 8;
    uint32_t* qu = reinterpret_cast<uint32_t*>(p);
    if (QU) p += size_t(blockDim.x) * 8;
    uint32_t* hk = reinterpret_cast<uint32_t*>(p);  // 8-byte aligned (pair reads)
    p += size_t(H) * 4;
    uint32_t* L = reinterpret_cast<uint32_t*>(p);
    p += size_t(Dmax) * 4;
    uint32_t* rp = reinterpret_cast<uint32_t*>(p);
    p += size_t(Dmax + 1) * 4;
    uint32_t* fb = reinterpret_cast<uint32_t*>(p);
    p += size_t(FW) * 4;
    uint16_t* hv = reinterpret_cast<uint16_t*>(p);
    __shared__ uint32_t sTask, sTmp[16], sSpan, sTile;
    const RootSet rs{hk, hv, H - 1, hshift, fb, fshift};
    unsigned long long acc = 0;
    uint32_t qcur = 0xFFFFFFFFu;
    const uint64_t* rowptr = nullptr;
    const uint32_t* col = nullptr;
    TaskChunk chunk(ntasks, gridDim.x, 8);
    while (true) {
        if (threadIdx.x == 0) {
            sTask = thread_next_task(chunk, cursor);
            sTile = 0;
        }
        __syncthreads();
        const uint32_t t = wave_bcast(sTask);
        if (t >= ntasks) break;
        const uint32_t q = task_q ? task_q[t] : 0u;
        if (q != qcur) {
            if (qcur != 0xFFFFFFFFu) flush_query(acc, qcur, results);
            qcur = q;
            rowptr = queries[q].rowptr;
            col = queries[q].col;
        }
        const uint32_t u = task_v[t];
        const uint64_t ub = ldg(rowptr, u);
        const uint32_t d = wave_bcast(uint32_t(ldg(rowptr, u + 1) - ub));
        for (uint32_t i = threadIdx.x; i < d; i += blockDim.x) L[i] = ldg(col, ub + i);
        root_clear(rs, H, FW, threadIdx.x, blockDim.x);
        if (K >= 4)
            for (uint32_t i = threadIdx.x, n = TRI ? tri_words(d) : d * W; i < n; i += blockDim.x) rows[i] = 0;
        __syncthreads();
        for (uint32_t i = threadIdx.x; i < d; i += blockDim.x) root_insert(rs, L[i], uint16_t(i));
        __syncthreads();
        build_rows_block<K, NG, QU, TRI>(rowptr, col, L, d, rs, 0, d, rows, d, rb, rp, sTmp, &sSpan, acc, qu);
        if (K == 4) count_rows_mfma4<TRI, (WPE <= 4)>(rows, d, d, &sTile, acc);
        else if (K == 5) {
            if (W <= 8) count5_edges<TRI>(rows, d, rp, sTmp, &sTile, acc);
            else count_rows<K, true, TRI>(rows, d, d, 0, 0, acc, &sTile);
        }
        __syncthreads();
    }
    if (qcur != 0xFFFFFFFFu) flush_query(acc, qcur, results);
}

// ------------------------------------------------------------ hub tasks --
// Hub roots are split: item c of hub h builds rows [c*RPC, (c+1)*RPC) in an
// LDS buffer and stores them to the hub's HBM slab (k >= 4), or counts the
// matches directly (k = 3).  Dynamic LDS: [row buffer RPC*Wmax u64 (k>=4)]
// [row descriptors RPC u64][probe queues 64 (id, row) per wave (QU)][hash
// keys H u32][L Dmax u32][group prefix RPC+1 u32][filter FW u32][hash vals
// H u16].
constexpr uint32_t kHubItemRun = 8;

template <int K, bool QU = false>
__global__ __launch_bounds__(1024) void hub_rows_kernel(const QueryRef* __restrict__ queries,
                                                       const HubMeta* __restrict__ hubs, uint32_t nhubs,
                                                       const uint32_t* __restrict__ item_prefix, uint32_t nitems,
                                                       uint32_t* __restrict__ cursor,
                                                       unsigned long long* __restrict__ results,
                                                       unsigned long long* __restrict__ slab, uint32_t Dmax,
                                                       uint32_t Wmax, uint32_t RPC, uint32_t H, uint32_t FW,
                                                       uint32_t fkey) {
    extern __shared__ __align__(16) unsigned char smem[];
    unsigned char* p = smem;
    unsigned long long* rbuf = nullptr;
    if (K >= 4) {
        rbuf = reinterpret_cast<unsigned long long*>(p);
        p += size_t(RPC) * Wmax * 8;
    }
    unsigned long long* rb = reinterpret_cast<unsigned long long*>(p);
    p += size_t(RPC) * 8;
    uint32_t* qu = reinterpret_cast<uint32_t*>(p);
    if (QU) p += size_t(blockDim.x) * 8;
    uint32_t* hk = reinterpret_cast<uint32_t*>(p);  // 8-byte aligned (pair reads)
    p += size_t(H) * 4;
    uint32_t* L = reinterpret_cast<uint32_t*>(p);
    p += size_t(Dmax) * 4;
    uint32_t* rp = reinterpret_cast<uint32_t*>(p);
    p += size_t(RPC + 1) * 4;
    uint32_t* fb = reinterpret_cast<uint32_t*>(p);
    p += size_t(FW) * 4;
    uint16_t* hv = reinterpret_cast<uint16_t*>(p);
    __shared__ uint32_t sItem, sTmp[16], sSpan;
    unsigned long long acc = 0;
    uint32_t qcur = 0xFFFFFFFFu;
    uint32_t hcur = 0xFFFFFFFFu;
    const uint64_t* rowptr = nullptr;
    const uint32_t* col = nullptr;
    RootSet rs{hk, hv, 0, 0, fb, 0};
    // items come in runs of at least kHubItemRun: consecutive items are
    // mostly the same hub's, whose list and hash are staged once per run
    TaskChunk chunk(nitems, gridDim.x, 8, kHubItemRun);
    while (true) {
        if (threadIdx.x == 0) sItem = thread_next_task(chunk, cursor);
        __syncthreads();
        const uint32_t it = sItem;
        if (it >= nitems) break;
        const uint32_t h = row_of(item_prefix, nhubs, it);
        const HubMeta hm = hubs[h];
        if (hm.q != qcur) {
            if (qcur != 0xFFFFFFFFu) flush_query(acc, qcur, results);
            qcur = hm.q;
            rowptr = queries[hm.q].rowptr;
            col = queries[hm.q].col;
        }
        const uint32_t d = hm.d;
        if (h != hcur) {  // stage this hub's list and hash once per run of its items
            uint32_t Hh = 64;
            while (Hh < 2 * d) Hh <<= 1;
            uint32_t bits = 0;
            while ((1u << bits) < Hh) ++bits;
            uint32_t fw = 0;
            rs.hmask = Hh - 1;
            rs.hshift = 32 - bits;
            filter_geometry(d, &fw, &rs.fshift, fkey);
            const uint64_t ub = ldg(rowptr, hm.u);
            for (uint32_t i = threadIdx.x; i < d; i += blockDim.x) L[i] = ldg(col, ub + i);
            root_clear(rs, Hh, fw, threadIdx.x, blockDim.x);
            __syncthreads();
            for (uint32_t i = threadIdx.x; i < d; i += blockDim.x) root_insert(rs, L[i], uint16_t(i));
            hcur = h;
        }
        const uint32_t r0 = (it - item_prefix[h]) * RPC;
        const uint32_t r1 = r0 + RPC < d ? r0 + RPC : d;
        if (K >= 4)
            for (uint32_t i = threadIdx.x; i < (r1 - r0) * hm.W; i += blockDim.x) rbuf[i] = 0;
        __syncthreads();
        build_rows_block<K, kGroups, QU>(rowptr, col, L, d, rs, r0, r1, rbuf, r1 - r0, rb, rp, sTmp, &sSpan, acc, qu);
        if (K >= 4) {
            // the slab is word-major over all d rows of the hub
            const uint32_t nr = r1 - r0;
            unsigned long long* dst = slab + hm.slab_off + r0;
            for (uint32_t i = threadIdx.x; i < nr * hm.W; i += blockDim.x) {
                const uint32_t kk = i / nr;
                dst[size_t(kk) * d + (i - kk * nr)] = rbuf[i];
            }
        }
        __syncthreads();
    }
    if (qcur != 0xFFFFFFFFu) flush_query(acc, qcur, results);
}

// One wave per hub row: counts the cliques whose second vertex is that row.
template <int K>
__global__ __launch_bounds__(256) void hub_count_kernel(const HubMeta* __restrict__ hubs, uint32_t nhubs,
                                                        const uint32_t* __restrict__ row_prefix, uint32_t nrows,
                                                        uint32_t* __restrict__ cursor,
                                                        unsigned long long* __restrict__ results,
                                                        const unsigned long long* __restrict__ slab) {
    unsigned long long acc = 0;
    uint32_t qcur = 0xFFFFFFFFu;
    TaskChunk chunk(nrows, gridDim.x * (blockDim.x >> 6), 8);
    while (true) {
        const uint32_t r = wave_next_task(chunk, cursor);
        if (r >= nrows) break;
        const uint32_t h = row_of(row_prefix, nhubs, r);
        const HubMeta hm = hubs[h];
        if (hm.q != qcur) {
            if (qcur != 0xFFFFFFFFu) flush_query(acc, qcur, results);
            qcur = hm.q;
        }
        const uint32_t i = r - row_prefix[h];
        count_rows<K, false>(slab + hm.slab_off, hm.d, hm.d, i, hm.d, acc);
    }
    if (qcur != 0xFFFFFFFFu) flush_query(acc, qcur, results);
}

// 4-cliques of the hubs on MFMA: one wave per 2 x 2 block of 32 x 32 tiles
// of a hub's A A^T (mfma4_block), the blocks of every hub (tile_prefix:
// exclusive prefix over hubs of W (W + 1) / 2, W = ceil(d / 64)) taken from
// one cursor, read from the hub's word-major slab rows in HBM.
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void hub_mfma4_kernel(const HubMeta* __restrict__ hubs, uint32_t nhubs,
                                                        const uint32_t* __restrict__ tile_prefix, uint32_t ntiles,
                                                        uint32_t* __restrict__ cursor,
                                                        unsigned long long* __restrict__ results,
                                                        const unsigned long long* __restrict__ slab) {
    unsigned long long acc = 0;
    uint32_t qcur = 0xFFFFFFFFu;
    TaskChunk chunk(ntiles, gridDim.x * (blockDim.x >> 6), 8);
    while (true) {
        const uint32_t t = wave_next_task(chunk, cursor);
        if (t >= ntiles) break;
        const uint32_t h = wave_bcast(row_of(tile_prefix, nhubs, t));
        const HubMeta hm = hubs[h];
        if (hm.q != qcur) {
            if (qcur != 0xFFFFFFFFu) flush_query(acc, qcur, results);
            qcur = hm.q;
        }
        uint32_t I, J;
        tile_coords(t - tile_prefix[h], &I, &J);
        acc += mfma4_block<false>(slab + hm.slab_off, hm.d, hm.d, I, J);
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
    // measured best on power-law graphs (RMAT scale 22): the LDS-heavy
    // buckets in 1024-thread workgroups, bucket kernels one after another
    const uint32_t hi[kNumBuckets] = {64, 128, 512, 1024, 0xFFFFFFFFu};
    const uint32_t blk[kNumBuckets] = {256, 256, 512, 1024, 1024};
    for (int b = 0; b < kNumBuckets; ++b) {
        c.hi[b] = hi[b];
        c.block[b] = blk[b];
    }
    c.serial = 1;
    return c;
}

int check_buckets(const BucketCfg& c) {
    if (c.hi[0] < 2 || c.hi[0] > 64) KZ_FAIL(KZ_ERR_INVALID, "first bucket bound must be in 2..64");
    for (int b = 1; b < kNumBuckets - 1; ++b)
        if (c.hi[b] <= c.hi[b - 1]) KZ_FAIL(KZ_ERR_INVALID, "bucket bounds must increase");
    if (c.hi[kNumBuckets - 2] > 1024) KZ_FAIL(KZ_ERR_INVALID, "LDS buckets end at 1024 or below");
    for (int b = 0; b < kNumBuckets; ++b)
        if (c.block[b] != 256 && c.block[b] != 512 && c.block[b] != 1024)
            KZ_FAIL(KZ_ERR_INVALID, "block sizes must be 256, 512 or 1024");
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
    for (auto& h : hub_meta) hipFree(h.second);
    for (auto& h : hub_items) hipFree(h.second);
    for (auto& h : hub_rows) hipFree(h.second);
    for (auto& h : hub_tiles) hipFree(h.second);
}

// Must match the carving at the top of clique_block_kernel / hub_rows_kernel.
static uint32_t filter_words(uint32_t Dmax, uint32_t per_key = 16) {
    uint32_t w, sh;
    filter_geometry(Dmax, &w, &sh, per_key);
    return w;
}

// queue_threads: the workgroup size when it keeps probe queues, else 0;
// tri: packed rows (k >= 4).
static size_t block_lds_bytes(int K, uint32_t Dmax, uint32_t W, uint32_t H, uint32_t queue_threads = 0,
                              bool tri = false) {
    const size_t rows = K < 4 ? 0 : tri ? size_t(tri_words(Dmax)) : size_t(Dmax) * W;
    return rows * 8 + size_t(Dmax) * 8 + size_t(queue_threads) * 8 + size_t(Dmax) * 4 +
           size_t(Dmax + 1) * 4 + size_t(filter_words(Dmax)) * 4 + size_t(H) * 6 + 16;
}

static void hash_size(uint32_t Dmax, uint32_t* H, uint32_t* hshift);

// Largest degree in (lo, hi] whose 4-clique rows, packed, with probe queues
// fit two 512-thread workgroups per CU while full rows fit one; 0 if none.
static uint32_t two_per_cu_degree(uint32_t lo, uint32_t hi) {
    const size_t per_block = rt().lds_per_block;
    auto per_cu = [&](size_t b) { return per_block / (b + kStaticLdsReserve); };
    for (uint32_t d = hi; d > lo; --d) {
        uint32_t H, hshift;
        hash_size(d, &H, &hshift);
        const uint32_t W = (d + 63) / 64;
        const size_t ldst = block_lds_bytes(4, d, W, H, 512, true);
        if (ldst + kStaticLdsReserve <= per_block && per_cu(ldst) >= 2)
            return per_cu(block_lds_bytes(4, d, W, H)) == 1 ? d : 0;
    }
    return 0;
}

static size_t hub_lds_bytes(int K, uint32_t Dmax, uint32_t Wmax, uint32_t RPC, uint32_t H,
                            uint32_t queue_threads = 0, uint32_t fkey = 16) {
    return (K >= 4 ? size_t(RPC) * Wmax * 8 : 0) + size_t(RPC) * 8 + size_t(queue_threads) * 8 +
           size_t(RPC + 1) * 4 + size_t(Dmax) * 4 + size_t(filter_words(Dmax, fkey)) * 4 + size_t(H) * 6 + 16;
}

// Launch shape of the hub rows kernel for row chunks of rpc rows: LDS bytes,
// workgroup size and workgroups per CU (by LDS, and by its 80 VGPRs: 24
// waves per CU).  1024-thread workgroups run one per CU; where the LDS
// leaves room for three 512-thread ones, those fill the CU instead (RMAT-24
// 4-cliques 211.2 -> 205.3 ms; where it does not, as for the forward
// degrees of RMAT-26/ef32, 1024 threads stay faster: 6.42 against 7.95 s).
constexpr uint32_t kHubWavesPerCU = 24;
struct HubShape {
    size_t lds;
    uint32_t block, per_cu;
};
static HubShape hub_shape(int K, uint32_t dmax, uint32_t wmax, uint32_t rpc, uint32_t H, uint32_t cfg_block,
                          bool queues, uint32_t fkey = 16) {
    const size_t per_block = rt().lds_per_block;
    auto fit = [&](size_t b) { return std::max<uint32_t>(1, uint32_t(per_block / (b + kStaticLdsReserve))); };
    HubShape s{hub_lds_bytes(K, dmax, wmax, rpc, H, queues ? cfg_block : 0, fkey), cfg_block, 0};
    s.per_cu = fit(s.lds);
    if (cfg_block == 1024) {
        const size_t l512 = hub_lds_bytes(K, dmax, wmax, rpc, H, queues ? 512 : 0, fkey);
        if (fit(l512) >= 3) s = HubShape{l512, 512, fit(l512)};
    }
    s.per_cu = std::min<uint32_t>(s.per_cu, std::max<uint32_t>(1, kHubWavesPerCU / (s.block / 64)));
    return s;
}

static int set_lds_attrs() {
    static bool done = false;
    if (done) return KZ_OK;
    // Dynamic LDS may use the block limit minus the kernels' static LDS; the
    // attribute is advisory on this runtime, so a refusal is not fatal (the
    // launch itself reports a size it cannot honour).
    const int maxb = int(rt().lds_per_block) - kStaticLdsReserve;
    const void* fns[] = {reinterpret_cast<const void*>(&clique_block_kernel<3>),
                         reinterpret_cast<const void*>(&clique_block_kernel<4>),
                         reinterpret_cast<const void*>(&clique_block_kernel<5>),
                         reinterpret_cast<const void*>(&clique_block_kernel<4, 3, 4>),
                         reinterpret_cast<const void*>(&clique_block_kernel<4, 2, 7>),
                         reinterpret_cast<const void*>(&clique_block_kernel<3, 2, 7>),
                         reinterpret_cast<const void*>(&clique_block_kernel<3, kGroups, 6, true>),
                         reinterpret_cast<const void*>(&clique_block_kernel<4, kGroups, 6, true>),
                         reinterpret_cast<const void*>(&clique_block_kernel<5, kGroups, 6, true>),
                         reinterpret_cast<const void*>(&clique_block_kernel<4, 3, 4, true>),
                         reinterpret_cast<const void*>(&clique_block_kernel<4, 2, 7, true>),
                         reinterpret_cast<const void*>(&clique_block_kernel<3, 2, 7, true>),
                         reinterpret_cast<const void*>(&clique_block_kernel<4, kGroups, 6, true, true>),
                         reinterpret_cast<const void*>(&clique_block_kernel<5, kGroups, 6, true, true>),
                         reinterpret_cast<const void*>(&clique_block_kernel<5, 3, 4>),
                         reinterpret_cast<const void*>(&clique_block_kernel<5, 3, 4, true>),
                         reinterpret_cast<const void*>(&clique_block_kernel<5, 3, 4, true, true>),
                         reinterpret_cast<const void*>(&clique_block_kernel<4, 3, 4, true, true>),
                         reinterpret_cast<const void*>(&clique_block_kernel<4, 4, 4, true, true>),
                         reinterpret_cast<const void*>(&clique_block_kernel<4, 2, 7, true, true>),
                         reinterpret_cast<const void*>(&hub_rows_kernel<3>),
                         reinterpret_cast<const void*>(&hub_rows_kernel<4>),
                         reinterpret_cast<const void*>(&hub_rows_kernel<5>),
                         reinterpret_cast<const void*>(&hub_rows_kernel<4, true>),
                         reinterpret_cast<const void*>(&hub_rows_kernel<5, true>)};
    for (const void* f : fns)
        if (hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, maxb) != hipSuccess)
            (void)hipGetLastError();
    done = true;
    return KZ_OK;
}

static void hash_size(uint32_t Dmax, uint32_t* H, uint32_t* hshift) {
    uint32_t h = 64;
    while (h < 2 * Dmax) h <<= 1;
    uint32_t bits = 0;
    while ((1u << bits) < h) ++bits;
    *H = h;
    *hshift = 32 - bits;
}

// Arguments of clique_block_kernel after the launch configuration.
struct BlockArgs {
    const QueryRef* qrefs;
    const uint32_t* tv;
    const uint32_t* tq;
    uint32_t nt;
    uint32_t* cur;
    unsigned long long* result;
    uint32_t Dmax, W, H, hshift, FW, fshift;
};

template <int K, int NG, int WPE>
static void launch_block(bool qu, bool tri, uint32_t grid, uint32_t block, size_t lds, hipStream_t s,
                         const BlockArgs& a) {
    if constexpr (K >= 4) {
        if (tri) {  // packed rows, with queues (where full rows and queues do not fit)
            hipLaunchKernelGGL((clique_block_kernel<K, NG, WPE, true, true>), dim3(grid), dim3(block), lds, s, a.qrefs,
                               a.tv, a.tq, a.nt, a.cur, a.result, a.Dmax, a.W, a.H, a.hshift, a.FW, a.fshift);
            return;
        }
    }
    if (qu)
        hipLaunchKernelGGL((clique_block_kernel<K, NG, WPE, true>), dim3(grid), dim3(block), lds, s, a.qrefs, a.tv,
                           a.tq, a.nt, a.cur, a.result, a.Dmax, a.W, a.H, a.hshift, a.FW, a.fshift);
    else
        hipLaunchKernelGGL((clique_block_kernel<K, NG, WPE, false>), dim3(grid), dim3(block), lds, s, a.qrefs, a.tv,
                           a.tq, a.nt, a.cur, a.result, a.Dmax, a.W, a.H, a.hshift, a.FW, a.fshift);
}

// KZ_STREAM_QUEUE=0 turns the block buckets' probe queues off.
static bool stream_queue_enabled() {
    static const bool on = [] {
        const char* e = std::getenv("KZ_STREAM_QUEUE");
        return !(e && e[0] == '0');
    }();
    return on;
}

// One (k, bucket) launch (two kernels for hubs).
template <int K>
static int launch_group(CliqueJob& x, const CliqueJob::Group& gr, hipStream_t s) {
    const Runtime& R = rt();
    const uint32_t nt = uint32_t(gr.count);
    const QueryRef* qrefs = static_cast<const QueryRef*>(x.d_queries);
    const uint32_t* tv = x.d_task_v + gr.first;
    const uint32_t* tq = x.d_task_q ? x.d_task_q + gr.first : nullptr;
    uint32_t* cur = x.d_cursor + gr.cursor;
    const size_t lds_limit = R.lds_per_block - kStaticLdsReserve;
    if (gr.bucket == 0) {
        const uint32_t grid = std::min<uint32_t>((nt + 3) / 4, uint32_t(R.num_cus) * 8);
        if (stream_queue_enabled())
            hipLaunchKernelGGL((clique_wave_kernel<K, true>), dim3(grid), dim3(256), 0, s, qrefs, tv, tq, nt, cur,
                               x.d_result);
        else
            hipLaunchKernelGGL((clique_wave_kernel<K, false>), dim3(grid), dim3(256), 0, s, qrefs, tv, tq, nt, cur,
                               x.d_result);
        return KZ_OK;
    }
    if (gr.bucket < kNumBuckets - 1) {
        const uint32_t Dmax = gr.dmax;
        const uint32_t W = (Dmax + 63) / 64;
        uint32_t H, hshift;
        hash_size(Dmax, &H, &hshift);
        const uint32_t block = gr.block ? gr.block : x.cfg.block[gr.bucket];
        const size_t lds = block_lds_bytes(K, Dmax, W, H);
        if (lds > lds_limit)
            KZ_FAIL(KZ_ERR_UNSUPPORTED, "bucket %d needs %zu bytes of LDS (limit %zu)", gr.bucket, lds, lds_limit);
        auto groups_per_cu = [&](size_t b) {
            const uint32_t n = std::max<uint32_t>(1, uint32_t(R.lds_per_block / (b + kStaticLdsReserve)));
            return std::min<uint32_t>(n, 2048 / block);
        };
        uint32_t per_cu = groups_per_cu(lds);
        // Probe queues where the LDS holds them at the same occupancy; for
        // 4-cliques whose full rows leave no room for them (the 512-1024
        // bucket), packed rows and queues.  Packed rows alone cost 0.4-2%
        // per bucket (RMAT-24: 128-512 72.09 -> 72.37 ms, 64-128 8.50 ->
        // 8.68), queues in the 512-1024 bucket gain 11% (66.09 -> 58.21).
        // 512-thread groups whose full rows hold them to one per CU take
        // packed rows where those fit two (4 waves per SIMD).  5-cliques,
        // whose count dominates, take packed rows wherever they keep the
        // occupancy (constant word offsets; RMAT-22 128-512 bucket 40.28 ->
        // 39.13 ms).
        const size_t ldsq = block_lds_bytes(K, Dmax, W, H, block);
        const size_t ldst = block_lds_bytes(K, Dmax, W, H, block, true);
        const bool qu = stream_queue_enabled() && ldsq <= lds_limit && groups_per_cu(ldsq) == per_cu;
        const bool tri_fit = K >= 4 && stream_queue_enabled() && ldst <= lds_limit;
        const bool tri2 = tri_fit && block == 512 && per_cu == 1 && groups_per_cu(ldst) >= 2;
        const bool tri = tri2 || (tri_fit && (K == 5 || !qu) && groups_per_cu(ldst) >= per_cu);
        if (tri2) per_cu = 2;
        const size_t lds_run = tri ? ldst : qu ? ldsq : lds;
        const uint32_t grid = std::min<uint32_t>(nt, uint32_t(R.num_cus) * per_cu);
        uint32_t FW, fshift;
        filter_geometry(Dmax, &FW, &fshift);
        const BlockArgs a{qrefs, tv, tq, nt, cur, x.d_result, Dmax, W, H, hshift, FW, fshift};
        if constexpr (K == 3) {  // no bitmap: the LDS leaves 7 waves per SIMD or more
            if (per_cu * block / 64 >= 28) {
                launch_block<3, 2, 7>(qu, false, grid, block, lds_run, s, a);
                return KZ_OK;
            }
        }
        if constexpr (K == 4) {
            if (per_cu * block / 64 >= 28) {  // LDS leaves 7 waves per SIMD or more
                launch_block<4, 2, 7>(qu, tri, grid, block, lds_run, s, a);
                return KZ_OK;
            }
            if (per_cu * block / 64 <= 16) {  // at most 4 waves per SIMD
                // with packed rows and queues, 4 groups per lane fit in 128
                // VGPRs without spills (RMAT-24 4-cliques 188.6 -> 186.1 ms)
                if (tri) launch_block<4, 4, 4>(qu, tri, grid, block, lds_run, s, a);
                else launch_block<4, 3, 4>(qu, tri, grid, block, lds_run, s, a);
                return KZ_OK;
            }
        }
        if constexpr (K == 5) {
            if (per_cu * block / 64 <= 16) {  // at most 4 waves per SIMD: 128 VGPRs
                launch_block<5, 3, 4>(qu, tri, grid, block, lds_run, s, a);
                return KZ_OK;
            }
        }
        launch_block<K, kGroups, 6>(qu, tri, grid, block, lds_run, s, a);
        return KZ_OK;
    }
    // hubs: rows kernel, then (k >= 4) the count kernel on the same stream
    const CliqueJob::HubPlan& hp = x.hubplan[K];
    uint32_t H, hshift;
    hash_size(hp.dmax, &H, &hshift);
    (void)hshift;
    const HubShape sh = hub_shape(K, hp.dmax, hp.wmax, hp.rpc, H, x.cfg.block[gr.bucket], hp.queues, hp.fkey);
    if (sh.lds > lds_limit)
        KZ_FAIL(KZ_ERR_UNSUPPORTED, "hub rows need %zu bytes of LDS (limit %zu)", sh.lds, lds_limit);
    const uint32_t grid = std::min<uint32_t>(hp.nitems, uint32_t(R.num_cus) * sh.per_cu);
    const HubMeta* meta = static_cast<const HubMeta*>(x.hub_meta[K]);
    if (hp.queues)
        hipLaunchKernelGGL((hub_rows_kernel<K, true>), dim3(grid), dim3(sh.block), sh.lds, s, qrefs, meta, hp.nhubs,
                           x.hub_items[K], hp.nitems, cur, x.d_result, K >= 4 ? x.slabs[K] : nullptr, hp.dmax,
                           hp.wmax, hp.rpc, H, filter_words(hp.dmax, hp.fkey), hp.fkey);
    else
        hipLaunchKernelGGL((hub_rows_kernel<K, false>), dim3(grid), dim3(sh.block), sh.lds, s, qrefs, meta, hp.nhubs,
                           x.hub_items[K], hp.nitems, cur, x.d_result, K >= 4 ? x.slabs[K] : nullptr, hp.dmax,
                           hp.wmax, hp.rpc, H, filter_words(hp.dmax, hp.fkey), hp.fkey);
    if (K == 4) {
        const uint32_t cgrid = std::min<uint32_t>((hp.ntiles + 3) / 4, uint32_t(R.num_cus) * 8);
        hipLaunchKernelGGL(hub_mfma4_kernel, dim3(cgrid), dim3(256), 0, s, meta, hp.nhubs, x.hub_tiles[K], hp.ntiles,
                           cur + kCursorStride, x.d_result, x.slabs[K]);
    } else if (K >= 5) {
        const uint32_t cgrid = std::min<uint32_t>((hp.nrows + 3) / 4, uint32_t(R.num_cus) * 8);
        hipLaunchKernelGGL(hub_count_kernel<K>, dim3(cgrid), dim3(256), 0, s, meta, hp.nhubs, x.hub_rows[K], hp.nrows,
                           cur + kCursorStride, x.d_result, x.slabs[K]);
    }
    return KZ_OK;
}

// Hub plan for one k: per-hub metadata, item and row prefixes, HBM slab.
static int prepare_hubs(CliqueJob& x, int k, const std::vector<uint32_t>& hv, const std::vector<uint32_t>& hq,
                        const std::vector<uint32_t>& hd) {
    CliqueJob::HubPlan hp;
    hp.nhubs = uint32_t(hv.size());
    hp.dmax = 0;
    hp.wmax = 0;
    for (uint32_t d : hd) {
        hp.dmax = std::max(hp.dmax, d);
        hp.wmax = std::max(hp.wmax, (d + 63) / 64);
    }
    uint32_t H, hshift;
    hash_size(hp.dmax, &H, &hshift);
    const size_t lds_limit = rt().lds_per_block - kStaticLdsReserve;
    hp.rpc = 64;
    while (hp.rpc > 1 && hub_lds_bytes(k, hp.dmax, hp.wmax, hp.rpc, H) > lds_limit) hp.rpc >>= 1;
    // probe queues (k >= 4: the instances within 80 VGPRs) where the LDS
    // holds them beside the same chunk with as many threads per CU.  Not by
    // halving the chunk: rows of 32 instead of 64 with queues measured
    // slower (RMAT-24 hub rows 17.53 -> 18.83 ms, RMAT-26/ef32 2.93 ->
    // 3.12 s, there 16 instead of 32).  Where the rows kernel runs one
    // 1024-thread workgroup per CU, an 8-bit-per-id filter makes room for
    // them (RMAT-26/ef32 hub rows 2,919 -> 2,718 ms); where it runs three
    // 512-thread ones, that filter costs more than the queues gain (RMAT-24
    // 17.45 -> 18.21 ms).
    if (k >= 4 && stream_queue_enabled()) {
        const uint32_t cfg_block = x.cfg.block[kNumBuckets - 1];
        const HubShape s0 = hub_shape(k, hp.dmax, hp.wmax, hp.rpc, H, cfg_block, false);
        for (uint32_t fkey : {16u, 8u}) {
            if (fkey == 8 && !(s0.block == 1024 && s0.per_cu == 1)) break;
            const HubShape sq = hub_shape(k, hp.dmax, hp.wmax, hp.rpc, H, cfg_block, true, fkey);
            if (sq.lds <= lds_limit && sq.per_cu * sq.block >= s0.per_cu * s0.block) {
                hp.queues = true;
                hp.fkey = fkey;
                break;
            }
        }
    }
    std::vector<HubMeta> meta(hp.nhubs);
    std::vector<uint32_t> items(hp.nhubs), rows(hp.nhubs), tiles(hp.nhubs);
    uint64_t slab = 0, nt = 0;
    uint32_t ni = 0, nr = 0;
    for (uint32_t h = 0; h < hp.nhubs; ++h) {
        const uint32_t d = hd[h];
        const uint32_t W = (d + 63) / 64;
        const uint64_t nb = (d + 63) / 64;  // 2 x 2 tile blocks: one per pair of 64-bit words
        meta[h] = HubMeta{hv[h], hq[h], d, W, slab};
        slab += uint64_t(d) * W;
        items[h] = ni;
        rows[h] = nr;
        tiles[h] = uint32_t(nt);
        ni += (d + hp.rpc - 1) / hp.rpc;
        nr += d;
        nt += nb * (nb + 1) / 2;
    }
    if (nt >= (uint64_t(1) << 32)) KZ_FAIL(KZ_ERR_UNSUPPORTED, "%llu hub tiles (> 2^32)", (unsigned long long)nt);
    hp.nitems = ni;
    hp.nrows = nr;
    hp.ntiles = uint32_t(nt);
    void* dm = nullptr;
    uint32_t *di = nullptr, *dr = nullptr, *dt = nullptr;
    KZ_HIP(hipMalloc(&dm, sizeof(HubMeta) * std::max<uint32_t>(1, hp.nhubs)));
    KZ_HIP(hipMalloc(&di, 4 * std::max<uint32_t>(1, hp.nhubs)));
    KZ_HIP(hipMalloc(&dr, 4 * std::max<uint32_t>(1, hp.nhubs)));
    KZ_HIP(hipMalloc(&dt, 4 * std::max<uint32_t>(1, hp.nhubs)));
    x.bytes += (sizeof(HubMeta) + 12) * std::max<uint32_t>(1, hp.nhubs);
    x.hub_meta[k] = dm;
    x.hub_items[k] = di;
    x.hub_rows[k] = dr;
    x.hub_tiles[k] = dt;
    KZ_HIP(hipMemcpy(dm, meta.data(), sizeof(HubMeta) * meta.size(), hipMemcpyHostToDevice));
    KZ_HIP(hipMemcpy(di, items.data(), 4 * items.size(), hipMemcpyHostToDevice));
    KZ_HIP(hipMemcpy(dr, rows.data(), 4 * rows.size(), hipMemcpyHostToDevice));
    KZ_HIP(hipMemcpy(dt, tiles.data(), 4 * tiles.size(), hipMemcpyHostToDevice));
    if (k >= 4) {
        unsigned long long* s = nullptr;
        KZ_HIP(hipMalloc(&s, std::max<uint64_t>(1, slab) * 8));
        x.bytes += std::max<uint64_t>(1, slab) * 8;
        x.slabs[k] = s;
    }
    x.hubplan[k] = hp;
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
    // Tasks grouped per (k, bucket), each group ordered by degree descending
    // and, within a degree, by query and vertex: a counting sort on the
    // degree (<= kMaxBitmapDegree), two passes over the roots.
    constexpr int kLists = 3 * kNumBuckets;
    const uint32_t D = kMaxBitmapDegree;
    std::vector<std::vector<uint64_t>> cnt(kLists);
    auto list_of = [&](int k, uint32_t d) { return (k - 3) * kNumBuckets + int(bucket_of(cfg, d)); };
    for (uint32_t qi = 0; qi < queries.size(); ++qi) {
        const CliqueQuery& Q = queries[qi];
        if (Q.k < 3 || Q.k > 5) KZ_FAIL(KZ_ERR_UNSUPPORTED, "clique size %d outside 3..5", Q.k);
        const std::vector<uint32_t>& fd = *Q.fdeg;
        const uint32_t vend = uint32_t(std::min<uint64_t>(fd.size(), Q.root_hi));
        for (uint32_t v = Q.root_lo; v < vend; ++v) {
            const uint32_t d = fd[v];
            if (d < uint32_t(Q.k - 1)) continue;
            if (d > D)
                KZ_FAIL(KZ_ERR_UNSUPPORTED,
                        "persistent clique form needs out-degree <= %u (found %u); orient the graph by degree", D, d);
            auto& c = cnt[list_of(Q.k, d)];
            if (c.empty()) c.assign(size_t(D) + 1, 0);
            c[d]++;
        }
    }
    // list starts, then per degree (descending) the next free position
    std::vector<uint64_t> lstart(kLists + 1, 0);
    for (int l = 0; l < kLists; ++l) {
        uint64_t tot = 0;
        for (uint64_t c : cnt[l]) tot += c;
        lstart[l + 1] = lstart[l] + tot;
    }
    std::vector<std::vector<uint64_t>> pos(kLists);
    for (int l = 0; l < kLists; ++l) {
        if (cnt[l].empty()) continue;
        pos[l].assign(size_t(D) + 1, 0);
        uint64_t run = lstart[l];
        for (uint32_t d = D + 1; d-- > 0;) {
            pos[l][d] = run;
            run += cnt[l][d];
        }
    }
    const uint64_t ntask = lstart[kLists];
    std::vector<uint32_t> hv(ntask), hq(ntask);
    for (uint32_t qi = 0; qi < queries.size(); ++qi) {
        const CliqueQuery& Q = queries[qi];
        const std::vector<uint32_t>& fd = *Q.fdeg;
        const uint32_t vend = uint32_t(std::min<uint64_t>(fd.size(), Q.root_hi));
        for (uint32_t v = Q.root_lo; v < vend; ++v) {
            const uint32_t d = fd[v];
            if (d < uint32_t(Q.k - 1)) continue;
            const uint64_t at = pos[list_of(Q.k, d)][d]++;
            hv[at] = v;
            hq[at] = qi;
        }
    }
    x.groups.clear();
    int cursor = 0;
    for (int l = 0; l < kLists; ++l) {
        if (lstart[l + 1] == lstart[l]) continue;
        CliqueJob::Group gr;
        gr.k = 3 + l / kNumBuckets;
        gr.bucket = l % kNumBuckets;
        gr.first = lstart[l];
        gr.count = lstart[l + 1] - lstart[l];
        const uint32_t dfirst = (*queries[hq[gr.first]].fdeg)[hv[gr.first]];  // the largest in the group
        gr.dmax = gr.bucket == 0 ? 64 : (gr.bucket == kNumBuckets - 1 ? dfirst : cfg.hi[gr.bucket]);
        // 4-cliques, last bitmap bucket: the roots whose packed rows and
        // probe queues fit two 512-thread workgroups per CU (degree <= the
        // split) run as a group of their own, so one root's barriers and
        // count overlap another root's stream on the same CU.  (5-cliques,
        // whose count dominates, measured slower split: RMAT-22 113.8 ->
        // 124.6 ms in that bucket.)
        if (gr.k == 4 && gr.bucket == kNumBuckets - 2 && stream_queue_enabled()) {
            const uint32_t split = two_per_cu_degree(cfg.hi[gr.bucket - 1], gr.dmax);
            const uint64_t end = gr.first + gr.count;
            uint64_t at = gr.first;  // tasks run in descending degree
            while (split && at < end && (*queries[hq[at]].fdeg)[hv[at]] > split) ++at;
            if (split && at < end) {
                if (at > gr.first) {
                    CliqueJob::Group hi = gr;
                    hi.count = at - gr.first;
                    hi.cursor = cursor;
                    cursor += int(kCursorStride);
                    x.groups.push_back(hi);
                }
                gr.first = at;
                gr.count = end - at;
                gr.dmax = split;
                gr.block = 512;
            }
        }
        gr.cursor = cursor;
        // one cursor line per kernel (hubs: rows and count kernels)
        cursor += int(kCursorStride) * ((gr.bucket == kNumBuckets - 1) ? 2 : 1);
        if (gr.bucket == kNumBuckets - 1) {
            std::vector<uint32_t> v(hv.begin() + long(gr.first), hv.begin() + long(gr.first + gr.count));
            std::vector<uint32_t> q(hq.begin() + long(gr.first), hq.begin() + long(gr.first + gr.count));
            std::vector<uint32_t> d(gr.count);
            for (uint64_t i = 0; i < gr.count; ++i) d[i] = (*queries[q[i]].fdeg)[v[i]];
            if (int e = prepare_hubs(x, gr.k, v, q, d)) return e;
        }
        x.groups.push_back(gr);
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
    x.bytes += sizeof(QueryRef) * qr.size() + 4 * std::max<size_t>(1, hv.size()) * (queries.size() > 1 ? 2 : 1) +
               4 * size_t(std::max(1, cursor)) + 8 * std::max<size_t>(1, queries.size());
    // capture on a stream of its own: other threads may use the library
    // stream meanwhile (the server prepares jobs beside running queries)
    hipStream_t main = nullptr;
    KZ_HIP(hipStreamCreateWithFlags(&main, hipStreamNonBlocking));
    struct StreamFree {
        hipStream_t s;
        ~StreamFree() { (void)hipStreamDestroy(s); }
    } main_free{main};
    const size_t nbr = std::max<size_t>(1, x.groups.size());
    std::vector<hipStream_t> br(nbr, nullptr);
    std::vector<hipEvent_t> join(nbr, nullptr);
    hipEvent_t fork;
    for (size_t i = 0; i < nbr; ++i) {
        KZ_HIP(hipStreamCreateWithFlags(&br[i], hipStreamNonBlocking));
        KZ_HIP(hipEventCreateWithFlags(&join[i], hipEventDisableTiming));
    }
    KZ_HIP(hipEventCreateWithFlags(&fork, hipEventDisableTiming));
    // the launches' errors are read after the capture: start from none
    (void)hipGetLastError();
    KZ_HIP(hipStreamBeginCapture(main, hipStreamCaptureModeThreadLocal));
    hipLaunchKernelGGL(reset_u32_kernel, dim3(1), dim3(256), 0, main, x.d_cursor, std::max(1, cursor));
    hipLaunchKernelGGL(reset_u64_kernel, dim3(1), dim3(256), 0, main, x.d_result,
                       int(std::max<size_t>(1, queries.size())));
    KZ_HIP(hipEventRecord(fork, main));
    int e = KZ_OK;
    // serial: one chain on the main stream; otherwise one branch per group
    for (size_t i = 0; i < x.groups.size() && e == KZ_OK; ++i) {
        hipStream_t s = cfg.serial ? main : br[i];
        if (!cfg.serial) KZ_HIP(hipStreamWaitEvent(s, fork, 0));
        const CliqueJob::Group& gr = x.groups[i];
        if (gr.k == 3) e = launch_group<3>(x, gr, s);
        else if (gr.k == 4) e = launch_group<4>(x, gr, s);
        else e = launch_group<5>(x, gr, s);
    }
    if (!cfg.serial)
        for (size_t i = 0; i < x.groups.size(); ++i) {
            KZ_HIP(hipEventRecord(join[i], br[i]));
            KZ_HIP(hipStreamWaitEvent(main, join[i], 0));
        }
    hipGraph_t graph;
    KZ_HIP(hipStreamEndCapture(main, &graph));
    if (e == KZ_OK) {
        const hipError_t le = hipGetLastError();
        if (le != hipSuccess) {
            e = KZ_ERR_DEVICE;
            kz::set_error("a bucket launch was rejected during capture: %s", hipGetErrorString(le));
        }
    }
    for (size_t i = 0; i < nbr; ++i) {
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

int clique_job_launch(CliqueJob& x, hipStream_t s, unsigned long long* host_results) {
    KZ_HIP(hipGraphLaunch(x.exec, s));
    KZ_HIP(hipMemcpyAsync(host_results, x.d_result, sizeof(unsigned long long) * std::max<uint32_t>(1, x.nqueries),
                          hipMemcpyDeviceToHost, s));
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
    const uint32_t bounds[][4] = {{64, 128, 512, 1024}, {64, 192, 384, 1024}, {32, 256, 512, 1024}, {64, 256, 512, 1024}};
    const uint32_t blocks[][kNumBuckets] = {{256, 256, 256, 512, 512}, {256, 256, 512, 1024, 1024}};
    for (const auto& bd : bounds)
        for (const auto& bk : blocks)
            for (uint32_t serial = 0; serial < 2; ++serial) {
                BucketCfg c = default_buckets();
                for (int b = 0; b < 4; ++b) c.hi[b] = bd[b];
                for (int b = 0; b < kNumBuckets; ++b) c.block[b] = bk[b];
                c.serial = serial;
                out.push_back(c);
            }
    return out;
}

}  // namespace kz
