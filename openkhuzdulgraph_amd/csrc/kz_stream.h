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
    return wave_lane(x, 63u);
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
    return wave_bcast(total);
}

// Four consecutive list entries, read with one 16-byte load.
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

// Groups of 4 entries each lane keeps in flight per streaming iteration.
constexpr int kGroups = 2;

// Descriptor of one row, the neighbour list N+(v) = col[vb, ve), with
// `cand` candidates left after it.  A streamed row keeps vb in bits 0..60 and
// ve & 3 in bits 62..63; it is read as aligned 16-byte groups from vb & ~3
// (*groups of them), the first group starting at entry vb & 3 and the last
// one ending at ve & 3 (0: full).  A skewed row, whose list is more than
// kSkewFactor times longer than its candidates, is not streamed
// (*groups = 0): its candidates are binary-searched in the list; bit 61 marks
// it, bits 0..39 hold vb and bits 40..60 the length (lists too long or too
// far into col for those fields are streamed).  *len gets ve - vb.  Returns
// whether the row is skewed.  Column arrays are allocated with
// col_alloc_bytes, so the last group stays inside the allocation.
constexpr uint64_t kSkewBit = 1ull << 61;
constexpr uint64_t kStartMask = kSkewBit - 1;
constexpr int kSkewLenShift = 40;
constexpr uint64_t kSkewStartMask = (1ull << kSkewLenShift) - 1;
constexpr uint64_t kSkewMaxLen = (1ull << (61 - kSkewLenShift)) - 1;
constexpr uint64_t kSkewFactor = 4;

__device__ __forceinline__ bool row_desc(const uint64_t* rowptr, uint32_t v, uint32_t cand, uint64_t* desc,
                                         uint32_t* groups, uint32_t* len) {
    const uint64_t vb = ldg(rowptr, v), ve = ldg(rowptr, v + 1);
    *len = uint32_t(ve - vb);
    if (ve - vb > kSkewFactor * cand && ve - vb <= kSkewMaxLen && vb <= kSkewStartMask) {
        *desc = vb | ((ve - vb) << kSkewLenShift) | kSkewBit;
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
// The hash probes in pairs: a key's first two slots are its home slot h and
// h ^ 1, read together with one 8-byte load (hk is 8-byte aligned, hv
// 4-byte aligned so that a pair's two indices are one 4-byte load), then
// (h | 1) + 1, ... linearly.
struct RootSet {
    uint32_t* hk;
    uint16_t* hv;
    uint32_t hmask, hshift;
    uint32_t* fb;
    uint32_t fshift;
};

// Filter of a root set: 2^(32 - fshift) bits in 32-bit words (fewer where
// the bits below the hash slot run out).  An id's word is the bits of its
// Fibonacci product (hash_slot) just below the slot's, its bit within the
// word the product's low 5 bits (the id's low 5 bits under an odd
// multiplier, a bijection), so one 32-bit multiply serves slot, word and
// bit.  The word's bits start at filter_offset + 5 of the product.
__device__ __forceinline__ uint32_t filter_offset(uint32_t hshift, uint32_t fshift) {
    const uint32_t fbits = 32 - fshift;
    return fbits < hshift ? hshift - fbits : 0u;
}
__device__ __forceinline__ uint32_t filter_word_of(uint32_t prod, uint32_t hshift, uint32_t fshift) {
    const uint32_t off = filter_offset(hshift, fshift);
    return __builtin_amdgcn_ubfe(prod, off + 5, hshift - off - 5);
}
// The id's bit of its filter word: v_bfe_u32 reads its offset modulo 32,
// so the product goes in as it is.
__device__ __forceinline__ uint32_t filter_test(uint32_t word, uint32_t prod) {
    return __builtin_amdgcn_ubfe(word, prod, 1u);
}

// Clears the hash (H slots) and filter (fw words) over threads tid..T.
__device__ __forceinline__ void root_clear(const RootSet& s, uint32_t H, uint32_t fw, uint32_t tid, uint32_t T) {
    for (uint32_t i = tid; i < H; i += T) s.hk[i] = kEmpty;
    for (uint32_t i = tid; i < fw; i += T) s.fb[i] = 0;
}

// Inserts x at the first free slot of its probe order (h, h ^ 1, then
// (h | 1) + 1 onwards); concurrent inserts of distinct keys are safe.
__device__ __forceinline__ void root_insert(const RootSet& s, uint32_t x, uint16_t idx) {
    const uint32_t h0 = hash_slot(x, s.hshift);
    uint32_t h = h0;
    if (atomicCAS(&s.hk[h], kEmpty, x) != kEmpty) {
        h = h0 ^ 1u;
        if (atomicCAS(&s.hk[h], kEmpty, x) != kEmpty) {
            h = ((h0 | 1u) + 1u) & s.hmask;
            while (atomicCAS(&s.hk[h], kEmpty, x) != kEmpty) h = (h + 1) & s.hmask;
        }
    }
    s.hv[h] = idx;
    const uint32_t prod = x * 2654435761u;
    atomicOr(&s.fb[filter_word_of(prod, s.hshift, s.fshift)], 1u << (prod & 31u));
}

// Filter geometry for lists of up to dmax ids: words and shift.
// per_key: filter bits per id of the list (16 by default).
__host__ __device__ inline void filter_geometry(uint32_t dmax, uint32_t* words, uint32_t* fshift,
                                                uint32_t per_key = 16) {
    uint32_t bits = 6;
    while ((1u << bits) < per_key * dmax && bits < 31) ++bits;
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

// Largest i in [lo, hi) with p[i] <= f, given p[lo] <= f (p non-decreasing);
// wave-uniform arguments and result.  Each round reads 64 evenly spaced
// entries, one per lane, and narrows the range 64-fold with one ballot, so
// ranges up to 4096 take two LDS round trips.
__device__ __forceinline__ uint32_t wave_search(const uint32_t* p, uint32_t lo, uint32_t hi, uint32_t f) {
    const uint32_t lane = threadIdx.x & 63;
    while (hi - lo > 1) {
        const uint32_t step = (hi - lo + 63) >> 6;
        const uint32_t i = lo + __umul24(lane, step);
        // lanes past hi read p[hi - 1] (no branch around the read)
        const unsigned long long b = __ballot((i < hi) & (p[i < hi ? i : hi - 1] <= f));
        lo += (uint32_t(__popcll(b)) - 1) * step;
        hi = hi < lo + step ? hi : lo + step;
    }
    return lo;
}

// Streams groups [0, G) of rows [0, nr) described by desc and gp (exclusive
// prefix of the group counts, nr + 1 entries) over threads tid = 0 .. T-1
// (whole waves).  Every entry x found in the root's set at index j
// is a hit: hit(r, j, p) with p the entry's position in col (j is only
// looked up when kIndex).
//
// A wave walks NG * 64 groups per iteration (a span), lane l taking groups
// l, l + 64, ...: the span's last row comes from one wave_search, and each
// lane finds the rows of its NG groups with NG binary searches run in
// lockstep.  After the NG 16-byte loads, the 4 * NG entries go through the
// root set in batched phases, each issuing all its LDS reads before it
// waits: filter words, then the first two hash slots of the entries that
// pass the filter, then the indices of the keys found.  Only keys displaced
// further (rare at a load factor of at most 1/2) probe on.
//
// Spans go to waves either as one contiguous share per wave (ctr == NULL;
// a span's first row is the previous span's last), or, for workgroups,
// from a counter in LDS (ctr, zeroed by the caller before a barrier), one
// span at a time with the next one's atomic issued ahead: the waves then
// finish together instead of waiting at the caller's barrier for the one
// whose share probed most.
// Probes the n (wave-uniform, <= 64) queued candidates of a span, one per
// lane: entry i = (id, row) at q[2i], q[2i + 1].  hit(r, j, 0).
template <bool kIndex, class Hit>
__device__ __forceinline__ void drain_queue(const uint32_t* q, uint32_t n, const RootSet& rs, Hit& hit) {
    wave_lds_sync();
    const uint32_t lane = threadIdx.x & 63;
    const bool act = lane < n;
    const uint2 ent = reinterpret_cast<const uint2*>(q)[lane];  // entries past n are stale and ignored
    const uint32_t x = ent.x, r = ent.y;
    const uint32_t hp = act ? (x * 2654435761u) >> (rs.hshift + 1) : 0u;
    const uint2 kp = reinterpret_cast<const uint2*>(rs.hk)[hp];
    const uint32_t vp = kIndex ? reinterpret_cast<const uint32_t*>(rs.hv)[hp] : 0u;
    const bool f0 = kp.x == x, f1 = kp.y == x;
    if (act & (f0 | f1)) hit(r, f1 ? vp >> 16 : vp & 0xFFFFu, uint64_t(0));
    if (act & !f0 & !f1 & (kp.x != kEmpty) & (kp.y != kEmpty)) {
        const uint32_t j = hash_probe(rs.hk, rs.hv, rs.hmask, (2 * hp + 2) & rs.hmask, x);
        if (j != kNotFound) hit(r, j, uint64_t(0));
    }
}

template <int NG, bool kIndex, bool kQueue, class Hit>
__device__ __forceinline__ void stream_span(const uint32_t* col, const unsigned long long* desc, const uint32_t* gp,
                                            uint32_t g0, uint32_t g1, uint32_t rlo, uint32_t rhi, const RootSet& rs,
                                            uint32_t* q, Hit&& hit) {
    constexpr int NE = 4 * NG;
    const uint32_t lane = threadIdx.x & 63;
    // rows of this lane's groups (past g1: the group g1, masked below)
    uint32_t gg[NG], rr[NG];
#pragma unroll
    for (int k = 0; k < NG; ++k) {
        const uint32_t g = g0 + uint32_t(k) * 64 + lane;
        gg[k] = g < g1 ? g : g1;
        rr[k] = rlo;
    }
    for (uint32_t len = rhi - rlo; len > 1;) {
        const uint32_t half = len >> 1;
#pragma unroll
        for (int k = 0; k < NG; ++k) {
            const uint32_t m = rr[k] + half;
            rr[k] = gp[m] <= gg[k] ? m : rr[k];
        }
        len -= half;
    }
    u32x4 y[NG];
    uint32_t lo[NG], hi[NG];
    uint64_t at[NG];
#pragma unroll
    for (int k = 0; k < NG; ++k) {
        const uint32_t r = rr[k];
        const uint32_t p0 = gp[r], p1 = gp[r + 1];
        const uint64_t ds = desc[r];
        const uint64_t vb = ds & kStartMask;
        const uint32_t e3 = uint32_t(ds >> 62);
        const uint32_t q = gg[k] - p0;
        const bool live = g0 + uint32_t(k) * 64 + lane <= g1;
        lo[k] = !live ? 4u : q ? 0u : uint32_t(vb & 3);
        hi[k] = (gg[k] + 1 == p1 && e3) ? e3 : 4u;
        at[k] = (vb & ~uint64_t(3)) + 4ull * q;
        y[k] = ldg(reinterpret_cast<const u32x4*>(col + (vb & ~uint64_t(3))), q);
    }
    // filter (one multiply per id gives the filter bit and the hash slot;
    // the conditions are built without branches)
    const uint32_t foff = filter_offset(rs.hshift, rs.fshift);
    const uint32_t fwid = rs.hshift - foff;  // filter bits, at least 6
    uint32_t prod[NE], fw[NE];
#pragma unroll
    for (int e = 0; e < NE; ++e) {
        prod[e] = y[e >> 2][e & 3] * 2654435761u;
        fw[e] = rs.fb[__builtin_amdgcn_ubfe(prod[e], foff + 5, fwid - 5)];
    }
    // (no test against the root's largest id: its compare cost more than the
    // probes it saved, 30.93 -> 30.71 ms on RMAT-22 4-cliques)
    if constexpr (kQueue) {
        // queued probes: the span's filter passes (about one id in ten) are
        // packed into the wave's queue, (id, row) each, and probed 64 at a
        // time, so the pair reads, compares and hit stores run once per
        // candidate instead of once per id slot.  Each pass condition goes
        // straight into its ballot (kept as a lane mask, not a flag).
        // entries [lo, hi) of each group as a 4-bit mask (0 for dead groups)
        uint32_t vm[NG];
#pragma unroll
        for (int k = 0; k < NG; ++k) vm[k] = (0xFu >> (4u - hi[k])) & (0xFu << lo[k]);
        uint32_t n = 0;  // wave-uniform
#pragma unroll
        for (int e = 0; e < NE; ++e) {
            const bool pe = (filter_test(fw[e], prod[e]) & __builtin_amdgcn_ubfe(vm[e >> 2], uint32_t(e & 3), 1u)) != 0u;
            const unsigned long long m = __builtin_amdgcn_ballot_w64(pe);  // the compare's mask as is
            const uint32_t c = uint32_t(__popcll(m));
            if (n + c > 64) {
                drain_queue<kIndex>(q, n, rs, hit);
                n = 0;
            }
            if (pe) {
                const uint32_t rank =
                    __builtin_amdgcn_mbcnt_hi(uint32_t(m >> 32), __builtin_amdgcn_mbcnt_lo(uint32_t(m), 0u));
                uint32_t* slot = q + 2 * n + 2 * rank;
                slot[0] = y[e >> 2][e & 3];
                slot[1] = rr[e >> 2];
            }
            n += c;
        }
        if (n) drain_queue<kIndex>(q, n, rs, hit);
        return;
    }
    bool pass[NE];
#pragma unroll
    for (int e = 0; e < NE; ++e) {
        const uint32_t c = uint32_t(e & 3);
        pass[e] = (c >= lo[e >> 2]) & (c < hi[e >> 2]) & (filter_test(fw[e], prod[e]) != 0u);
    }
    // home pairs (h & ~1, h | 1): keys with one 8-byte read, their indices
    // with one 4-byte read beside it (lanes without a candidate read pair 0:
    // a broadcast)
    bool found[NE], cont[NE];
    uint32_t hp[NE], jj[NE];
#pragma unroll
    for (int e = 0; e < NE; ++e) {
        const uint32_t x = y[e >> 2][e & 3];
        hp[e] = pass[e] ? prod[e] >> (rs.hshift + 1) : 0u;
        const uint2 kp = reinterpret_cast<const uint2*>(rs.hk)[hp[e]];
        const uint32_t vp = kIndex ? reinterpret_cast<const uint32_t*>(rs.hv)[hp[e]] : 0u;
        const bool f0 = kp.x == x, f1 = kp.y == x;
        found[e] = pass[e] & (f0 | f1);
        cont[e] = pass[e] & !f0 & !f1 & (kp.x != kEmpty) & (kp.y != kEmpty);
        jj[e] = f1 ? vp >> 16 : vp & 0xFFFFu;
    }
    bool any_cont = false;
#pragma unroll
    for (int e = 0; e < NE; ++e) {
        if (found[e]) hit(rr[e >> 2], jj[e], at[e >> 2] + uint32_t(e & 3));
        any_cont |= cont[e];
    }
    if (any_cont) {
#pragma unroll
        for (int e = 0; e < NE; ++e) {
            if (cont[e]) {
                const uint32_t x = y[e >> 2][e & 3];
                const uint32_t j = hash_probe(rs.hk, rs.hv, rs.hmask, (2 * hp[e] + 2) & rs.hmask, x);
                if (j != kNotFound) hit(rr[e >> 2], j, at[e >> 2] + uint32_t(e & 3));
            }
        }
    }
}

// kQueue: q is this wave's queue of 64 (id, row) pairs (128 u32) in LDS for
// queued probes (hits then get p = 0).
template <int NG = kGroups, bool kIndex = true, bool kQueue = false, class Hit>
__device__ __forceinline__ void stream_groups(const uint32_t* col, const unsigned long long* desc, const uint32_t* gp,
                                              uint32_t nr, uint32_t G, uint32_t tid, uint32_t T, const RootSet& rs,
                                              Hit&& hit, uint32_t* ctr = nullptr, uint32_t* q = nullptr) {
    constexpr uint32_t SPAN = 64 * NG;
    const uint32_t lane = tid & 63;
    // the same in every lane: kept scalar
    nr = wave_bcast(nr);
    G = wave_bcast(G);
    if (ctr) {
        const uint32_t nspan = (G + SPAN - 1) / SPAN;
        uint32_t c = 0;
        if (lane == 0) c = atomicAdd(ctr, 1u);
        c = wave_bcast(c);
        while (c < nspan) {
            uint32_t cn = 0;
            if (lane == 0) cn = atomicAdd(ctr, 1u);
            const uint32_t g0 = c * SPAN, g1 = (g0 + SPAN < G ? g0 + SPAN : G) - 1;
            const uint32_t rlo = wave_search(gp, 0, nr, g0);
            const uint32_t rhi = wave_search(gp, rlo, nr, g1) + 1;
            stream_span<NG, kIndex, kQueue>(col, desc, gp, g0, g1, rlo, rhi, rs, q, hit);
            c = wave_bcast(cn);
        }
        return;
    }
    const uint32_t nw = T >> 6;
    const uint32_t per = ((G + nw - 1) / nw + 63) & ~63u;
    const uint32_t wb = (tid >> 6) * per;
    const uint32_t wbeg = wb < G ? wb : G;
    const uint32_t wend = wbeg + per < G ? wbeg + per : G;
    if (wbeg >= wend) return;
    uint32_t rlo = wave_search(gp, 0, nr, wbeg);
    for (uint32_t g0 = wbeg; g0 < wend; g0 += SPAN) {
        const uint32_t g1 = (g0 + SPAN < wend ? g0 + SPAN : wend) - 1;
        const uint32_t rhi = wave_search(gp, rlo, nr, g1) + 1;
        stream_span<NG, kIndex, kQueue>(col, desc, gp, g0, g1, rlo, rhi, rs, q, hit);
        rlo = rhi - 1;
    }
}

// Position of x in the sorted list a[0, n), or n when absent.
__device__ __forceinline__ uint32_t find_global(const uint32_t* __restrict__ a, uint32_t n, uint32_t x) {
    const uint32_t i = lower_bound_global(a, n, x);
    return (i < n && ldg(a, i) == x) ? i : n;
}

// Candidates of the skewed rows among rows [0, nr) (list indices r0 + r of
// L, d ids; candidates of row r are L[r0 + r + 1 .. d)), searched in the
// rows' lists.  Wave `wave` of nw takes rows r = wave, wave + nw, ..., 64 at a
// time.  A search costs about one vector-L1 lookup per distinct cache line
// it touches (which bounds this phase, not latency: more searches in flight
// per lane measured slower), so rows are searched two ways:
//   * rows with fewer than kSampledRow candidates: their (row, candidate)
//     pairs are spread over the lanes, one binary search each (about
//     log2(len) - 3 lines);
//   * the others one at a time: the lanes first read 64 evenly spaced
//     entries of the row's list (one load), each candidate finds its 1/64 of
//     the list by a shuffle search over those samples and binary-searches
//     only that part (64 lines per row, then about 1 - 2 per candidate).
// No workgroup barrier, so it overlaps other waves' streaming.  hit(r, j, p):
// candidate j found at position p of col.  Call with every lane of the wave.
constexpr uint32_t kSampledRow = 12;  // RMAT-22 4-cliques: 4 / 8 / 12 / 16 / 24: 31.4 / 30.9 / 30.9 / 31.0 / 31.3 ms

template <class Hit>
__device__ __forceinline__ void search_skewed(const uint32_t* col, const unsigned long long* desc, const uint32_t* L,
                                              uint32_t d, uint32_t r0, uint32_t nr, uint32_t wave, uint32_t nw,
                                              Hit&& hit) {
    const uint32_t lane = threadIdx.x & 63;
    // the same in every lane: kept scalar
    d = wave_bcast(d);
    r0 = wave_bcast(r0);
    nr = wave_bcast(nr);
    wave = wave_bcast(wave);
    nw = wave_bcast(nw);
    for (uint32_t rbase = wave; rbase < nr; rbase += 64 * nw) {
        const uint32_t r = rbase + lane * nw;
        uint32_t cnt = 0, n = 0;
        uint64_t b = 0;
        if (r < nr) {
            const uint64_t ds = desc[r];
            if (ds & kSkewBit) {
                b = ds & kSkewStartMask;
                n = uint32_t((ds & kStartMask) >> kSkewLenShift);
                cnt = d - 1 - (r0 + r);
            }
        }
        const bool sampled = cnt >= kSampledRow;
        // light rows: one (row, candidate) pair per lane
        uint32_t excl;
        const uint32_t P = wave_exclusive_scan(sampled ? 0u : cnt, &excl);
        for (uint32_t pbase = 0; pbase < P; pbase += 64) {
            const uint32_t p = pbase + lane;
            // owning lane: the largest l with excl(l) <= p
            uint32_t l = 0;
#pragma unroll
            for (uint32_t step = 32; step; step >>= 1) {
                const uint32_t e = __shfl(excl, int(l + step));
                if (e <= p) l += step;
            }
            const bool live = p < P;
            const uint32_t rr = __shfl(r, int(l));
            const uint64_t bb = __shfl(b, int(l));
            const uint32_t jj = r0 + rr + 1 + (p - __shfl(excl, int(l)));
            // shuffles run in every lane: a pair's owner may be a lane
            // without a pair of its own in this round
            const uint32_t nn = __shfl(n, int(l));
            const uint32_t x = live ? L[jj] : 0u;
            uint32_t len = live ? nn : 0u, at = 0;
            // largest i with a[i] <= x (a = col + bb); v tracks a[at]
            uint32_t v = ldg(col, bb);
            while (__ballot(len > 1)) {
                const uint32_t half = len >> 1;
                const uint32_t t = ldg(col, bb + at + half);
                if (half && t <= x) {
                    at += half;
                    v = t;
                }
                len -= half;
            }
            if (live && v == x) hit(rr, jj, bb + at);
        }
        // heavy rows: sample, then search one 1/64 of the list.  Sample s
        // = a[floor(s * n / 64)] (a the row's list; n < 2^21, so s * n fits
        // a 24-bit multiply).  With n < 64 samples repeat, but the largest s
        // with sample s <= x then has a non-empty part (its successor's
        // sample differs), and x below sample 0 is absent.  The
        // heavy rows' candidates go 64 per iteration as one sequence: an
        // iteration may take the end of row A in lanes [0, split) and the
        // start of the next row B in the others, each row's samples in a
        // register of its own, so a row's last partial iteration does not
        // leave lanes idle.
        unsigned long long hm = __ballot(sampled);
        auto next_row = [&](uint32_t& rr, uint32_t& nn, uint32_t& cc, uint64_t& bb, uint32_t& smp) {
            const uint32_t hl = uint32_t(__builtin_ctzll(hm));
            hm &= hm - 1;
            rr = wave_lane(r, hl);
            nn = wave_lane(n, hl);
            cc = wave_lane(cnt, hl);
            bb = wave_lane(b, hl);
            smp = ldg(col + bb, __umul24(lane, nn) >> 6);
        };
        uint32_t rA = 0, nA = 0, cA = 0, sA = 0, oA = 0;  // oA: row A's candidates done
        uint64_t bA = 0;
        bool more = hm != 0;
        if (more) next_row(rA, nA, cA, bA, sA);
        while (more) {
            const uint32_t remA = cA - oA;
            const bool mixed = remA < 64 && hm != 0;
            uint32_t rB = 0, nB = 0, cB = 0, sB = 0;
            uint64_t bB = 0;
            if (mixed) next_row(rB, nB, cB, bB, sB);
            const uint32_t split = remA < 64 ? remA : 64u;
            const bool inB = lane >= split;
            const uint32_t c = inB ? lane - split : oA + lane;
            const bool live = inB ? (mixed && c < cB) : true;
            const uint32_t rr = inB ? rB : rA, nn = inB ? nB : nA;
            const uint64_t bb = inB ? bB : bA;
            const uint32_t jj = r0 + rr + 1 + c;
            const uint32_t x = live ? L[jj] : 0u;
            // the part: the largest s with sample s <= x (s = 0 if none)
            uint32_t sp = 0;
#pragma unroll
            for (uint32_t step = 32; step; step >>= 1) {
                uint32_t e = __shfl(sA, int(sp + step));
                if (mixed) {
                    const uint32_t eB = __shfl(sB, int(sp + step));
                    e = inB ? eB : e;
                }
                if (e <= x) sp += step;
            }
            uint32_t at = __umul24(sp, nn) >> 6;
            const uint32_t hi = sp == 63 ? nn : __umul24(sp + 1, nn) >> 6;
            uint32_t len = live ? hi - at : 0u;
            uint32_t v = __shfl(sA, int(sp));
            if (mixed) {
                const uint32_t vB = __shfl(sB, int(sp));
                v = inB ? vB : v;
            }
            const uint32_t* a = col + bb;
            while (__ballot(len > 1)) {
                const uint32_t half = len >> 1;
                const uint32_t t = ldg(a, at + half);
                if (half && t <= x) {
                    at += half;
                    v = t;
                }
                len -= half;
            }
            if (live && v == x) hit(rr, jj, bb + at);
            // row A of the next iteration
            if (remA > 64) {
                oA += 64;
            } else if (mixed && cB > 64 - split) {
                rA = rB;
                nA = nB;
                cA = cB;
                bA = bB;
                sA = sB;
                oA = 64 - split;
            } else if (hm) {
                next_row(rA, nA, cA, bA, sA);
                oA = 0;
            } else {
                more = false;
            }
        }
    }
}

}  // namespace kz
