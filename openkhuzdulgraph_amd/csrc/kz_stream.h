// Row streaming shared by the clique and census kernels: a root u's list
// L = N+(u) is staged in LDS with a hash (id -> index), and the lists
// N+(L[i]) of its rows are read as one stream of aligned 16-byte groups;
// every entry that is in L is a hit (row i, candidate j).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "kz_kernels.h"

namespace kz {

__device__ __forceinline__ void flush_query(unsigned long long& acc, uint32_t q, unsigned long long* results) {
    const unsigned long long s = wave_sum_u64(acc);
    if ((threadIdx.x & 63) == 0 && s) atomicAdd(&results[q], s);
    acc = 0;
}

// Largest i in [0, n) with p[i] <= f (p non-decreasing, p[0] == 0 <= f).
__device__ __forceinline__ uint32_t row_of(const uint32_t* p, uint32_t n, uint32_t f) {
    uint32_t lo = 0, hi = n;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (p[mid] <= f) lo = mid;
        else hi = mid;
    }
    return lo;
}

// Exclusive scan of one u32 per lane across a wave; returns the total.
__device__ __forceinline__ uint32_t wave_exclusive_scan(uint32_t v, uint32_t* excl) {
    const uint32_t lane = threadIdx.x & 63;
    uint32_t x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(x, o);
        if (lane >= uint32_t(o)) x += y;
    }
    *excl = x - v;
    return __shfl(x, 63);
}

// Block-wide exclusive scan in place over a[0..n) (n <= 64 * blockDim);
// tmp holds one u32 per wave.  Returns the total.  Call with all threads.
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t* a, uint32_t n, uint32_t* tmp) {
    const uint32_t per = (n + blockDim.x - 1) / blockDim.x;
    const uint32_t b = threadIdx.x * per;
    const uint32_t e = b + per < n ? b + per : n;
    uint32_t s = 0;
    for (uint32_t i = b; i < e; ++i) s += a[i];
    uint32_t excl;
    const uint32_t wtot = wave_exclusive_scan(s, &excl);
    const uint32_t lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) tmp[w] = wtot;
    __syncthreads();
    uint32_t base = 0, total = 0;
    const uint32_t nw = blockDim.x >> 6;
    for (uint32_t i = 0; i < nw; ++i) {
        if (i < w) base += tmp[i];
        total += tmp[i];
    }
    uint32_t run = base + excl;
    for (uint32_t i = b; i < e; ++i) {
        const uint32_t t = a[i];
        a[i] = run;
        run += t;
    }
    __syncthreads();
    return total;
}

// Four consecutive list entries, read with one 16-byte load.
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

// Groups of 4 entries each lane keeps in flight per streaming iteration.
constexpr int kGroups = 4;

// Descriptor of one row, the neighbour list N+(v) = col[vb, ve), with
// `cand` candidates left after it: bits 0..60 hold vb, bit 61 marks a
// "skewed" row and bits 62..63 hold ve & 3.  A streamed row is read as
// aligned 16-byte groups from vb & ~3 (*groups of them); the first group
// starts at entry vb & 3 and the last one ends at ve & 3 (0: full).  A
// skewed row, whose list is more than kSkewFactor times longer than its
// candidates, is not streamed (*groups = 0): its candidates are
// binary-searched in the list.  *len gets ve - vb.  Returns whether the row
// is skewed.  Column arrays are allocated with col_alloc_bytes, so the last
// group stays inside the allocation.
constexpr uint64_t kSkewBit = 1ull << 61;
constexpr uint64_t kStartMask = kSkewBit - 1;
constexpr uint64_t kSkewFactor = 4;

__device__ __forceinline__ bool row_desc(const uint64_t* rowptr, uint32_t v, uint32_t cand, uint64_t* desc,
                                         uint32_t* groups, uint32_t* len) {
    const uint64_t vb = ldg(rowptr, v), ve = ldg(rowptr, v + 1);
    *len = uint32_t(ve - vb);
    if (ve - vb > kSkewFactor * cand) {
        *desc = vb | kSkewBit;
        *groups = 0;
        return true;
    }
    *desc = vb | ((ve & 3) << 62);
    *groups = ve > vb ? uint32_t((((ve + 3) & ~uint64_t(3)) - (vb & ~uint64_t(3))) >> 2) : 0u;
    return false;
}

// The root's list in LDS: an open-addressing hash (id -> index) and a
// one-hash membership filter of 2^(32 - fshift) bits (at least 16 per key)
// that rejects most non-members with a single LDS read before the probe.
struct RootSet {
    uint32_t* hk;
    uint16_t* hv;
    uint32_t hmask, hshift;
    uint32_t* fb;
    uint32_t fshift;
};

__device__ __forceinline__ uint32_t filter_bit(uint32_t x, uint32_t fshift) { return (x * 0x85EBCA6Bu) >> fshift; }

// Clears the hash (H slots) and filter (fw words) over threads tid..T.
__device__ __forceinline__ void root_clear(const RootSet& s, uint32_t H, uint32_t fw, uint32_t tid, uint32_t T) {
    for (uint32_t i = tid; i < H; i += T) s.hk[i] = kEmpty;
    for (uint32_t i = tid; i < fw; i += T) s.fb[i] = 0;
}

__device__ __forceinline__ void root_insert(const RootSet& s, uint32_t x, uint16_t idx) {
    hash_insert(s.hk, s.hv, s.hmask, s.hshift, x, idx);
    const uint32_t b = filter_bit(x, s.fshift);
    atomicOr(&s.fb[b >> 5], 1u << (b & 31));
}

__device__ __forceinline__ uint32_t root_find(const RootSet& s, uint32_t x) {
    const uint32_t b = filter_bit(x, s.fshift);
    if (!((s.fb[b >> 5] >> (b & 31)) & 1u)) return kNotFound;
    return hash_find(s.hk, s.hv, s.hmask, s.hshift, x);
}

// Filter geometry for lists of up to dmax ids: words and shift.
__host__ __device__ inline void filter_geometry(uint32_t dmax, uint32_t* words, uint32_t* fshift) {
    uint32_t bits = 6;
    while ((1u << bits) < 16u * dmax && bits < 31) ++bits;
    *words = (1u << bits) >> 5;
    *fshift = 32 - bits;
}

// Largest i in [lo, hi) with p[i] <= f, given p[lo] <= f.
__device__ __forceinline__ uint32_t row_in(const uint32_t* p, uint32_t lo, uint32_t hi, uint32_t f) {
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (p[mid] <= f) lo = mid;
        else hi = mid;
    }
    return lo;
}

// Streams groups [0, G) of rows [0, nr) described by desc and gp (exclusive
// prefix of the group counts, nr + 1 entries) over threads tid = 0 .. T-1
// (whole waves).  Every entry x <= lmax found in the root's set at index j
// is a hit: hit(r, j, p) with p the entry's position in col.  Each wave
// takes 64 * kGroups consecutive groups per iteration: two wave-uniform
// searches bound their rows, each lane searches only that window, and each
// lane issues kGroups independent 16-byte loads (consecutive across the
// wave) before it probes.
template <class Hit>
__device__ __forceinline__ void stream_groups(const uint32_t* col, const unsigned long long* desc, const uint32_t* gp,
                                              uint32_t nr, uint32_t G, uint32_t tid, uint32_t T, const RootSet& rs,
                                              uint32_t lmax, Hit&& hit) {
    const uint32_t lane = tid & 63;
    const uint32_t wofs = (tid - lane) * kGroups;
    for (uint32_t cbase = 0; cbase < G; cbase += kGroups * T) {
        const uint32_t g0 = cbase + wofs;
        if (g0 >= G) continue;
        const uint32_t g1 = (g0 + 64 * kGroups < G ? g0 + 64 * kGroups : G) - 1;
        const uint32_t rlo = row_of(gp, nr, g0);
        const uint32_t rhi = row_in(gp, rlo, nr, g1) + 1;
        u32x4 y[kGroups];
        uint32_t rr[kGroups], lo[kGroups], hi[kGroups];
        uint64_t at[kGroups];
#pragma unroll
        for (int k = 0; k < kGroups; ++k) {
            const uint32_t g = g0 + uint32_t(k) * 64 + lane;
            y[k] = u32x4{0u, 0u, 0u, 0u};
            rr[k] = 0;
            lo[k] = 4;
            hi[k] = 0;
            at[k] = 0;
            if (g <= g1) {
                const uint32_t r = row_in(gp, rlo, rhi, g);
                const uint32_t q = g - gp[r];
                const uint64_t ds = desc[r];
                const uint64_t vb = ds & kStartMask;
                const uint32_t e3 = uint32_t(ds >> 62);
                rr[k] = r;
                lo[k] = q ? 0u : uint32_t(vb & 3);
                hi[k] = (g + 1 == gp[r + 1] && e3) ? e3 : 4u;
                at[k] = (vb & ~uint64_t(3)) + 4ull * q;
                y[k] = ldg(reinterpret_cast<const u32x4*>(col + (vb & ~uint64_t(3))), q);
            }
        }
#pragma unroll
        for (int k = 0; k < kGroups; ++k) {
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                const uint32_t x = y[k][c];
                if (uint32_t(c) >= lo[k] && uint32_t(c) < hi[k] && x <= lmax) {
                    const uint32_t j = root_find(rs, x);
                    if (j != kNotFound) hit(rr[k], j, at[k] + uint32_t(c));
                }
            }
        }
    }
}

// Position of x in the sorted list a[0, n), or n when absent.
__device__ __forceinline__ uint32_t find_global(const uint32_t* __restrict__ a, uint32_t n, uint32_t x) {
    const uint32_t i = lower_bound_global(a, n, x);
    return (i < n && ldg(a, i) == x) ? i : n;
}

// Candidates of the skewed rows among rows [0, nr) (list indices r0 + r of
// L, d ids; candidates of row r are L[r0 + r + 1 .. d)), binary-searched in
// the rows' lists.  Wave `wave` of nw takes rows r = wave, wave + nw, ...,
// 64 at a time, and spreads their (row, candidate) pairs over its lanes, so
// the searches of 64 pairs are in flight together; no workgroup barrier, so
// it overlaps other waves' streaming.  hit(r, j, p): candidate j found at
// position p of col.  row_len(r), called by the lane that takes row r, gives
// the row's list length.  Call with every lane of the wave.
template <class Hit, class Len>
__device__ __forceinline__ void search_skewed(const uint64_t* rowptr, const uint32_t* col,
                                              const unsigned long long* desc, const uint32_t* L, uint32_t d,
                                              uint32_t r0, uint32_t nr, uint32_t wave, uint32_t nw, Hit&& hit,
                                              Len&& row_len) {
    const uint32_t lane = threadIdx.x & 63;
    for (uint32_t rbase = wave; rbase < nr; rbase += 64 * nw) {
        const uint32_t r = rbase + lane * nw;
        uint32_t cnt = 0, n = 0;
        uint64_t b = 0;
        if (r < nr) {
            const uint64_t ds = desc[r];
            if (ds & kSkewBit) {
                b = ds & kStartMask;
                cnt = d - 1 - (r0 + r);
                n = row_len(r);
            }
        }
        uint32_t excl;
        const uint32_t P = wave_exclusive_scan(cnt, &excl);
        for (uint32_t pbase = 0; pbase < P; pbase += 64) {
            const uint32_t p = pbase + lane;
            // owning lane: the largest l with excl(l) <= p
            uint32_t l = 0;
#pragma unroll
            for (uint32_t step = 32; step; step >>= 1) {
                const uint32_t e = __shfl(excl, int(l + step));
                if (e <= p) l += step;
            }
            const uint32_t rr = __shfl(r, int(l));
            const uint64_t bb = __shfl(b, int(l));
            const uint32_t nn = __shfl(n, int(l));
            const uint32_t e0 = __shfl(excl, int(l));
            if (p < P) {
                const uint32_t j = r0 + rr + 1 + (p - e0);
                const uint32_t pos = find_global(col + bb, nn, L[j]);
                if (pos < nn) hit(rr, j, bb + pos);
            }
        }
    }
}

}  // namespace kz
