// Plain level-by-level executor of the plan IR (the "plain GPU form").
//
// Reference behaviour restated for the GPU:
//   * EmbeddingExplorationEngine::extend_embeddings (src/engine.cc:914-1118)
//     extends every embedding of one size, writing children into the next
//     level's queue; when that queue or its buffers are full the engine dives
//     one level deeper before continuing (src/engine.cc:1047-1082).  Here a
//     level is processed in two passes over its embeddings: pass A sizes every
//     embedding's children and cached prefixes (and adds COUNT outputs);
//     after an exclusive scan the parents are cut into chunks whose children
//     and prefixes fit the per-level HBM budget, and for each chunk pass B
//     writes them and the next level runs before the following chunk;
//   * an embedding is a record {parent index, new vertex}; the matched
//     vertices are rebuilt by walking parents (CompactExtendableEmbedding,
//     include/engine.h:795-805 and src/engine.cc:1008-1021);
//   * cached objects (include/engine.h:592-613) become per-slot arrays of
//     {offset, length} into a level's prefix buffer;
//   * the set operations are those of src/vertex_set.cc:20-86, and COUNT /
//     SCATTER outputs are the apps' candidate loops.
// MI355X mapping: one wave64 per embedding (persistent waves pulling batches
// from an atomic cursor); the shorter operand of an intersection (or the
// subtrahend) is staged in an LDS hash and the other operand is streamed
// with coalesced loads; heavily skewed pairs binary-search the long list
// instead.  Results are compacted in order with ballot/popcount, so every
// prefix is sorted exactly like the reference's merge output.

#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <chrono>
#include <cstdlib>
#include <cstring>
#include <type_traits>

#include "kz_device.h"
#include "kz_kernels.h"
#include "kz_plan.h"
#include "kz_setops.h"

namespace kz {

uint64_t g_level_budget = uint64_t(4) << 30;  // memory_size_per_level default (4 GiB)


constexpr int kWavesPerBlock = 4;
constexpr uint32_t kBatch = 64;  // embeddings per cursor grab: one per lane (siblings stay together)
constexpr int kFlatUnroll = 4;   // stream chunks loaded together by the flat path

struct DevLevel {
    uint64_t n = 0;
    const uint32_t* parent = nullptr;
    const uint32_t* vertex = nullptr;
    uint32_t root_base = 0;
    const uint64_t* slot_off[KZ_MAX_OPS] = {};
    const uint32_t* slot_len[KZ_MAX_OPS] = {};
    const uint32_t* slot_data[KZ_MAX_OPS] = {};
    uint64_t slot_base[KZ_MAX_OPS] = {};
};

struct LevelArgs {
    DevLevel lv[KZ_MAX_PATTERN + 1];
    int s;
    int mode;  // 0 = pass A (sizes + counts), 1 = pass B (write)
    uint64_t begin, end;
    const uint64_t* rowptr;
    const uint32_t* col;
    const kz_step_list* lists;
    const uint8_t* fuse;   // per list: ops whose result only feeds COUNT outputs
    const uint8_t* types;  // type_table row for size s
    uint32_t* cursor;
    unsigned long long* agg;
    uint32_t* err;
    // pass A outputs (indexed e - begin)
    uint32_t* nchild;
    uint32_t* slot_len_out[KZ_MAX_OPS];
    // pass B outputs
    const uint64_t* child_off;  // exclusive prefix, indexed e - begin... absolute prefix
    uint64_t child_base;
    uint32_t* child_parent;
    uint32_t* child_vertex;
    const uint64_t* slot_off_out[KZ_MAX_OPS];
    uint64_t slot_base_out[KZ_MAX_OPS];
    uint32_t* slot_data_out[KZ_MAX_OPS];
    uint32_t cached_mask;
    uint32_t* temp;
    uint32_t temp_cap;
};

// Per-wave LDS of the flat path: each sibling lane's stream offset, list
// start and new vertex (kEmpty when it repeats an ancestor's vertex); with
// SIZES, per sibling its children and its cached results' sizes.
template <bool SIZES>
struct FlatLds {
    uint64_t start[64];
    uint32_t off[64];
    uint32_t w[64];
    uint64_t anc_idx[KZ_MAX_PATTERN];  // the segment's ancestors: indexes at sizes 1.. and vertices
    uint32_t anc_m[KZ_MAX_PATTERN];
    uint32_t sib[SIZES ? KZ_MAX_OPS + 1 : 1][64];
};

__device__ __forceinline__ void add_agg(unsigned long long* acc, int agg, unsigned long long v) {
    switch (agg) {
        case 0: acc[0] += v; break;
        case 1: acc[1] += v; break;
        case 2: acc[2] += v; break;
        case 3: acc[3] += v; break;
        case 4: acc[4] += v; break;
        case 5: acc[5] += v; break;
        case 6: acc[6] += v; break;
        default: acc[7] += v; break;
    }
}

__device__ __forceinline__ uint64_t readlane_u64(uint64_t v, uint32_t lane) {
    const uint32_t lo = __builtin_amdgcn_readlane(uint32_t(v), lane);
    const uint32_t hi = __builtin_amdgcn_readlane(uint32_t(v >> 32), lane);
    return (uint64_t(hi) << 32) | lo;
}

template <int N>
__device__ __forceinline__ void sort_small(uint32_t* v) {
#pragma unroll
    for (int i = 0; i < N; ++i)
#pragma unroll
        for (int j = 0; j + 1 < N - i; ++j) {
            const uint32_t lo = min(v[j], v[j + 1]), hi = max(v[j], v[j + 1]);
            v[j] = lo;
            v[j + 1] = hi;
        }
}

// Position class of x against sorted matched values z[0..N) that may
// repeat: 2 * (distinct values below x) + (x equals one of them).
template <int N>
__device__ __forceinline__ uint32_t pos_class_sorted(uint32_t x, const uint32_t* z) {
    uint32_t lt = 0;
    bool eq = false;
#pragma unroll
    for (int i = 0; i < N; ++i) {
        const bool distinct = i == 0 || z[i] != z[i - 1];
        lt += distinct && z[i] < x;
        eq |= z[i] == x;
    }
    return 2 * lt + eq;
}

// The same for the ancestors' sorted values plus one more vertex w (kEmpty
// when w repeats an ancestor: it then adds no class boundary).
template <int N>
__device__ __forceinline__ uint32_t pos_class_plus(uint32_t x, const uint32_t* z, uint32_t w) {
    const uint32_t c = pos_class_sorted<N>(x, z);
    return c + 2 * (w < x) + (w == x && !(c & 1u));
}

// Number of elements of the sorted list A whose position class against the
// sorted values z[0..N) is in pred: each class is a run of A between two
// binary-searched bounds (src/vertex_set.cc:20-86 restated as counts).
template <int N>
__device__ uint64_t count_classes(const uint32_t* A, uint32_t n, const uint32_t* z, uint32_t pred) {
    uint64_t c = 0;
    uint32_t from = 0, t = 0;
#pragma unroll
    for (int i = 0; i < N; ++i) {
        if (i > 0 && z[i] == z[i - 1]) continue;
        const uint32_t lo = from + lower_bound_global(A + from, n - from, z[i]);
        const uint32_t hi = lo + upper_bound_global(A + lo, n - lo, z[i]);
        if ((pred >> (2 * t)) & 1u) c += lo - from;
        if ((pred >> (2 * t + 1)) & 1u) c += hi - lo;
        from = hi;
        ++t;
    }
    if ((pred >> (2 * t)) & 1u) c += n - from;
    return c;
}

// Whether x is in the LDS hash, probing from slot h on.
__device__ __forceinline__ bool hash_find_from(const uint32_t* hk, uint32_t hmask, uint32_t h, uint32_t x) {
    while (true) {
        const uint32_t key = hk[h];
        if (key == x) return true;
        if (key == kEmpty) return false;
        h = (h + 1) & hmask;
    }
}

// Whether x is in the LDS hash: the first two probes branch-free (the
// table is sparse, so the loop rarely runs).
__device__ __forceinline__ bool hash_has2(const uint32_t* hk, uint32_t hmask, uint32_t hshift, uint32_t x) {
    const uint32_t h = (x * 2654435761u) >> hshift;
    const uint32_t k0 = hk[h], k1 = hk[(h + 1) & hmask];
    const bool found = k0 == x || (k0 != kEmpty && k1 == x);
    if (found || k0 == kEmpty || k1 == kEmpty) return found;
    return hash_find_from(hk, hmask, (h + 2) & hmask, x);
}

// Rebuild embedding e of size S: matched vertices m[0..S) and ancestor
// indexes idx[1..S] (idx[S] = e) by walking parents (src/engine.cc:1008-1021).
template <int S>
__device__ __forceinline__ void rebuild(const LevelArgs& a, uint64_t e, uint32_t* m, uint64_t* idx) {
    idx[S] = e;
    uint64_t cur = e;
#pragma unroll
    for (int t = S; t >= 2; --t) {
        m[t - 1] = a.lv[t].vertex[cur];
        cur = a.lv[t].parent[cur];
        idx[t - 1] = cur;
    }
    m[0] = a.lv[1].root_base + uint32_t(cur);
}

// One embedding extended by the whole wave: the general path for every
// step list.
__device__ void one_embedding(const LevelArgs& a, uint64_t e, uint32_t* hk, HashCache& hc, unsigned long long* acc,
                              uint32_t* temp) {
    const uint32_t lane = threadIdx.x & 63;
    const int s = a.s;
    // rebuild the embedding (src/engine.cc:1008-1021)
    uint32_t m[KZ_MAX_PATTERN];
    uint64_t idx[KZ_MAX_PATTERN + 1];
    idx[s] = e;
    uint64_t cur = e;
    for (int t = s; t >= 2; --t) {
        m[t - 1] = a.lv[t].vertex[cur];
        cur = a.lv[t].parent[cur];
        idx[t - 1] = cur;
    }
    m[0] = a.lv[1].root_base + uint32_t(cur);
    const uint8_t li = a.types[order_code(m, s)];
    if (li == 0xFF) {
        if (lane == 0) atomicOr(a.err, 1u);
        return;
    }
    const kz_step_list& L = a.lists[li];
    const uint32_t fuse = a.fuse[li];
    // sorted distinct matched values for position classes
    uint32_t z[KZ_MAX_PATTERN];
    int q = 0;
    for (int i = 0; i < s; ++i) {
        bool dup = false;
        for (int j = 0; j < q; ++j) dup |= z[j] == m[i];
        if (!dup) z[q++] = m[i];
    }
    for (int i = 1; i < q; ++i)
        for (int j = i; j > 0 && z[j - 1] > z[j]; --j) {
            const uint32_t t = z[j];
            z[j] = z[j - 1];
            z[j - 1] = t;
        }
    uint32_t rlen[KZ_MAX_OPS] = {0, 0, 0, 0};
    auto resolve = [&](const kz_set_ref& r) -> SetView {
        SetView v;
        if (r.kind == KZ_SET_NBR) {
            const uint32_t x = m[r.a];
            const uint64_t b = a.rowptr[x];
            v.p = a.col + b;
            v.n = uint32_t(a.rowptr[x + 1] - b);
        } else if (r.a == s) {
            v.p = temp + uint64_t(r.b) * a.temp_cap;
            v.n = rlen[r.b];
        } else {
            const DevLevel& Lv = a.lv[r.a];
            const uint64_t i = idx[r.a];
            v.p = Lv.slot_data[r.b] + (Lv.slot_off[r.b][i] - Lv.slot_base[r.b]);
            v.n = Lv.slot_len[r.b][i];
        }
        return v;
    };
    // an operand is shared by siblings when it comes from an ancestor
    auto shared = [&](const kz_set_ref& r) {
        return (r.kind == KZ_SET_NBR && int(r.a) < s - 1) || (r.kind == KZ_SET_SLOT && int(r.a) < s);
    };
    for (int j = 0; j < L.num_ops; ++j) {
        const kz_op& op = L.ops[j];
        if (fuse & (1u << j)) {
            // result only feeds COUNT outputs: count while streaming
            if (a.mode != 0) continue;
            unsigned long long c[KZ_MAX_OUTS] = {0, 0, 0, 0};
            wave_set_op(op.op, resolve(op.lhs), resolve(op.rhs), shared(op.lhs), shared(op.rhs), hk, hc,
                        [&](uint32_t x, bool keep, uint64_t, uint32_t) {
                            const uint32_t cls = keep ? pos_class(x, z, q) : 0;
                            for (int o = 0; o < L.num_outs; ++o) {
                                const kz_out& out = L.outs[o];
                                if (out.src.kind != KZ_SET_SLOT || out.src.a != s || out.src.b != j) continue;
                                c[o] += __popcll(__ballot(keep && ((out.pred >> cls) & 1u)));
                            }
                        });
            for (int o = 0; o < L.num_outs; ++o) {
                const kz_out& out = L.outs[o];
                if (out.src.kind == KZ_SET_SLOT && out.src.a == s && out.src.b == j)
                    add_agg(acc, out.aggregator, c[o]);
            }
            continue;
        }
        uint32_t* outp = temp + uint64_t(j) * a.temp_cap;
        rlen[j] = wave_set_op(op.op, resolve(op.lhs), resolve(op.rhs), shared(op.lhs), shared(op.rhs), hk, hc,
                              [&](uint32_t x, bool keep, uint64_t msk, uint32_t k) {
                                  if (keep) outp[k + __popcll(msk & lanemask_lt())] = x;
                              });
    }
    uint32_t nch = 0;
    uint64_t cbase = 0;
    if (a.mode == 1) cbase = a.child_off[e - a.begin] - a.child_base;
    for (int o = 0; o < L.num_outs; ++o) {
        const kz_out& out = L.outs[o];
        if (out.src.kind == KZ_SET_SLOT && out.src.a == s && (fuse & (1u << out.src.b))) continue;
        if (out.kind == KZ_OUT_COUNT && a.mode != 0) continue;
        const SetView src = resolve(out.src);
        uint32_t cnt = 0;
        for (uint32_t base = 0; base < src.n; base += 64) {
            const uint32_t i = base + lane;
            bool pass = false;
            uint32_t v = 0;
            if (i < src.n) {
                v = src.p[i];
                pass = (out.pred >> pos_class(v, z, q)) & 1u;
            }
            const uint64_t msk = __ballot(pass);
            if (out.kind == KZ_OUT_SCATTER && a.mode == 1 && pass) {
                const uint64_t at = cbase + nch + cnt + __popcll(msk & lanemask_lt());
                a.child_parent[at] = uint32_t(e);
                a.child_vertex[at] = v;
            }
            cnt += __popcll(msk);
        }
        if (out.kind == KZ_OUT_COUNT) add_agg(acc, out.aggregator, cnt);
        else nch += cnt;
    }
    if (a.mode == 0) {
        if (lane == 0) {
            if (a.nchild) a.nchild[e - a.begin] = nch;
            for (int j = 0; j < KZ_MAX_OPS; ++j)
                if (a.cached_mask & (1u << j)) a.slot_len_out[j][e - a.begin] = j < L.num_ops ? rlen[j] : 0;
        }
    } else {
        for (int j = 0; j < L.num_ops; ++j) {
            if (!(a.cached_mask & (1u << j))) continue;
            uint32_t* dst = a.slot_data_out[j] + (a.slot_off_out[j][e - a.begin] - a.slot_base_out[j]);
            const uint32_t* srcp = temp + uint64_t(j) * a.temp_cap;
            for (uint32_t i = lane; i < rlen[j]; i += 64) dst[i] = srcp[i];
        }
    }
}

// Flat path (pass A only).  Lanes [b, en) of the wave hold sibling
// embeddings -- children of one parent -- that take the same step list, and
// that list only counts: every op intersects or subtracts an ancestor's set
// P (the same for all ops) and the new vertex's list N(w), and every output
// is a COUNT of an op's result, of N(w) or of an ancestor's set.  The
// siblings' lists N(w) are concatenated and streamed by all 64 lanes at
// once (each lane finds its element's sibling by a binary search of the
// offsets), probing P's LDS hash once per element; counts of ancestor sets
// (and the P part of P \ N(w)) are class runs found by binary search.  The
// counts equal the per-embedding path's: the outputs only add up.
template <int S, bool SIZES>
__device__ void flat_segment(const LevelArgs& a, const kz_step_list& L, uint32_t b, uint32_t en, const uint32_t* m,
                             const uint64_t* idx, uint64_t wstart, uint32_t wlen, uint64_t e_lane, uint32_t* hk,
                             HashCache& hc, FlatLds<SIZES>& fl, unsigned long long* acc, unsigned long long* lacc) {
    static_assert(S >= 2, "the flat path needs an ancestor");
    const uint32_t lane = threadIdx.x & 63;
    const bool sib = lane >= b && lane < en;
    if constexpr (SIZES) {
#pragma unroll
        for (int r = 0; r < KZ_MAX_OPS + 1; ++r) fl.sib[r][lane] = 0;
    }
    // ancestors (the same on every lane of the segment), published by its
    // first lane for operand lookups by plan index
    wave_lds_sync();
    if (lane == b) {
#pragma unroll
        for (int t = 0; t < S - 1; ++t) fl.anc_m[t] = m[t];
#pragma unroll
        for (int t = 1; t < S; ++t) fl.anc_idx[t] = idx[t];
    }
    wave_lds_sync();
    uint32_t ms[S - 1];
#pragma unroll
    for (int t = 0; t < S - 1; ++t) ms[t] = fl.anc_m[t];
    auto resolve_shared = [&](const kz_set_ref& r) -> SetView {
        SetView v;
        if (r.kind == KZ_SET_NBR) {
            const uint32_t x = fl.anc_m[r.a];
            const uint64_t o = a.rowptr[x];
            v.p = a.col + o;
            v.n = uint32_t(a.rowptr[x + 1] - o);
        } else {
            const DevLevel& Lv = a.lv[r.a];
            const uint64_t i = fl.anc_idx[r.a];
            v.p = Lv.slot_data[r.b] + (Lv.slot_off[r.b][i] - Lv.slot_base[r.b]);
            v.n = Lv.slot_len[r.b][i];
        }
        return v;
    };
    auto varying = [](const kz_set_ref& r) { return r.kind == KZ_SET_NBR && int(r.a) == S - 1; };
    uint32_t zs[S - 1];
#pragma unroll
    for (int t = 0; t < S - 1; ++t) zs[t] = ms[t];
    sort_small<S - 1>(zs);
    const uint32_t w = m[S - 1];
    bool wdup = false;
#pragma unroll
    for (int t = 0; t < S - 1; ++t) wdup |= zs[t] == w;
    uint32_t zw[S];  // this lane's sibling: ancestors' values and w, sorted
#pragma unroll
    for (int t = 0; t < S - 1; ++t) zw[t] = zs[t];
    zw[S - 1] = w;
    sort_small<S>(zw);
    // per output: 0 = count an ancestor's set per sibling, 1 = stream and
    // keep P's members, 2 = stream and keep non-members, 3 = stream all;
    // bit 2: subtract the streamed count from a per-sibling count of P
    const bool has_ops = L.num_ops > 0;
    SetView P{nullptr, 0};
    if (has_ops) {
        const kz_op& o0 = L.ops[0];
        P = resolve_shared(varying(o0.lhs) ? o0.rhs : o0.lhs);
    }
    // Counters: one per output, and with SIZES one per cached op result
    // (its size).  Each goes to an aggregator (COUNT outputs: dest -1), to
    // the sibling's children (SCATTER outputs: dest 0) or to the sibling's
    // size of cached result j (dest 1 + j); -2: unused.
    constexpr int NC = SIZES ? KZ_MAX_OUTS + KZ_MAX_OPS : KZ_MAX_OUTS;
    auto op_mode = [&](const kz_op& op) -> uint32_t {
        if (op.op == KZ_OP_INTERSECT) return 1;
        return varying(op.lhs) ? 2u : (1u | 4u);
    };
    uint32_t omode[NC], opred[NC];
    int cdest[NC];
#pragma unroll
    for (int k = 0; k < NC; ++k) {
        omode[k] = 0;
        opred[k] = 0;
        cdest[k] = -2;
    }
    bool need_stream = false;
#pragma unroll
    for (int o = 0; o < KZ_MAX_OUTS; ++o) {
        if (o >= L.num_outs) break;
        const kz_out& out = L.outs[o];
        uint32_t md = 0;
        if (out.src.kind == KZ_SET_SLOT && out.src.a == S) md = op_mode(L.ops[out.src.b]);
        else if (varying(out.src)) md = 3;
        omode[o] = md;
        opred[o] = out.pred;
        cdest[o] = out.kind == KZ_OUT_COUNT ? -1 : 0;
        need_stream |= (md & 3) != 0;
    }
    if constexpr (SIZES) {
#pragma unroll
        for (int j = 0; j < KZ_MAX_OPS; ++j) {
            if (j >= L.num_ops || !(a.cached_mask & (1u << j))) continue;
            omode[KZ_MAX_OUTS + j] = op_mode(L.ops[j]);
            opred[KZ_MAX_OUTS + j] = ~0u;
            cdest[KZ_MAX_OUTS + j] = 1 + j;
            need_stream = true;
        }
    }
    if (need_stream) {
        const uint64_t start = sib ? wstart : 0;
        const uint32_t len = sib ? wlen : 0;
        uint32_t incl = len;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const uint32_t t = __shfl_up(incl, d);
            if (lane >= uint32_t(d)) incl += t;
        }
        const uint32_t T = __builtin_amdgcn_readlane(incl, 63);
        wave_lds_sync();
        fl.off[lane] = incl - len;
        fl.start[lane] = start;
        fl.w[lane] = wdup ? kEmpty : w;
        // P's hash: 16 slots per key where the table allows (probe runs of
        // one or two slots); only this path fills the table in the flat
        // kernels
        uint32_t hmask = 0, hshift = 0;
        const bool probe = has_ops && P.n > 0;
        const bool hashed = probe && P.n <= kHashCap / 2;
        if (hashed) {
            hash_geometry(min(8 * P.n, kHashCap / 2), &hmask, &hshift);
            if (!(hc.p == P.p && hc.n == P.n)) {
                for (uint32_t i = lane; i <= hmask; i += 64) hk[i] = kEmpty;
                wave_lds_sync();
                for (uint32_t i = lane; i < P.n; i += 64) {
                    const uint32_t x = P.p[i];
                    uint32_t h = (x * 2654435761u) >> hshift;
                    while (atomicCAS(&hk[h], kEmpty, x) != kEmpty) h = (h + 1) & hmask;
                }
                hc.p = P.p;
                hc.n = P.n;
            }
        }
        wave_lds_sync();
        // The streamed counters, compacted: the loop below is compiled for
        // 1, 2, 4 or NC of them (each costs a few VALU per position).
        int sidx[NC];
        int ns = 0;
#pragma unroll
        for (int k = 0; k < NC; ++k) {
            sidx[k] = -1;
            if (omode[k] & 3u) {
#pragma unroll
                for (int t = 0; t < NC; ++t)
                    if (t == ns) sidx[t] = k;
                ++ns;
            }
        }
        auto run_stream = [&](auto ns_tag) {
            constexpr int NS = decltype(ns_tag)::value;
            // per counter, the position classes counted for a member of P and
            // for a non-member, where it goes, and whether it subtracts
            uint32_t cls_in[NS], cls_out[NS], c[NS];  // c: this lane's counts
            int dst[NS];
            uint32_t neg[NS], agg[NS];
#pragma unroll
            for (int i = 0; i < NS; ++i) {
                cls_in[i] = cls_out[i] = c[i] = neg[i] = agg[i] = 0;
                dst[i] = -2;
#pragma unroll
                for (int k = 0; k < NC; ++k)
                    if (k == sidx[i]) {
                        const uint32_t md = omode[k] & 3u;
                        cls_in[i] = (md == 1 || md == 3) ? opred[k] : 0u;
                        cls_out[i] = (md == 2 || md == 3) ? opred[k] : 0u;
                        dst[i] = cdest[k];
                        neg[i] = omode[k] & 4u;
                        agg[i] = k < KZ_MAX_OUTS && k < L.num_outs ? L.outs[k].aggregator : 0u;
                    }
            }
            // SIZES: a lane's counts for its current sibling go to that
            // sibling's LDS counters when it moves on
            auto flush = [&](uint32_t sibling) {
                if constexpr (SIZES) {
#pragma unroll
                    for (int i = 0; i < NS; ++i)
                        if (dst[i] >= 0 && c[i]) {
                            atomicAdd(&fl.sib[dst[i]][sibling], neg[i] ? 0u - c[i] : c[i]);
                            c[i] = 0;
                        }
                }
            };
            if (hashed || !probe || SIZES) {
                // Each lane walks its own contiguous range of ceil(T / 64)
                // positions of the concatenation (its first sibling found once
                // by binary search), kFlatUnroll loads in flight, moving to
                // the next sibling when it reaches its offset: no per-position
                // search.
                const uint32_t q = (T + 63) / 64;
                uint32_t p = min(lane * q, T);
                const uint32_t e = min(p + q, T);
                uint32_t lo = b;  // the last sibling whose offset is <= p holds p (non-empty)
#pragma unroll
                for (uint32_t step = 32; step; step >>= 1) {
                    const uint32_t t = lo + step;
                    if (t < en && fl.off[t] <= p) lo = t;
                }
                uint32_t nxt = lo + 1 < en ? fl.off[lo + 1] : T;
                uint64_t adr = fl.start[lo] - fl.off[lo];  // col index of position p is adr + p (mod 2^64)
                uint32_t wv = fl.w[lo];
                // Each step takes up to kFlatUnroll positions of one sibling
                // (the loads clamped, not branched, to the sibling's end), so
                // a lane checks for the next sibling once per step.
                while (p < e) {
                    if (p == nxt) {
                        flush(lo);
                        do {
                            ++lo;
                            nxt = lo + 1 < en ? fl.off[lo + 1] : T;
                        } while (nxt == p);
                        adr = fl.start[lo] - fl.off[lo];
                        wv = fl.w[lo];
                    }
                    const uint32_t m = min(uint32_t(kFlatUnroll), min(e, nxt) - p);
                    uint32_t xs[kFlatUnroll];
#pragma unroll
                    for (int k = 0; k < kFlatUnroll; ++k) xs[k] = a.col[adr + p + (uint32_t(k) < m ? uint32_t(k) : 0u)];
#pragma unroll
                    for (int k = 0; k < kFlatUnroll; ++k) {
                        bool found = false;
                        if (probe)
                            found = hashed ? hash_has2(hk, hmask, hshift, xs[k]) : bsearch_global(P.p, P.n, xs[k]);
                        const uint32_t cls = pos_class_plus<S - 1>(xs[k], zs, wv);
                        const uint32_t fm = 0u - uint32_t(found);  // (a select of arrays would index them)
                        const uint32_t in = uint32_t(k) < m;
#pragma unroll
                        for (int i = 0; i < NS; ++i)
                            c[i] += in & ((((cls_in[i] & fm) | (cls_out[i] & ~fm)) >> cls) & 1u);
                    }
                    p += m;
                }
                flush(lo);
            } else {
                // P probed by binary search in HBM: 64 consecutive positions
                // per chunk, so the lanes' searches (values of one sorted
                // list) follow nearly the same path; kFlatUnroll chunks loaded
                // together, each position's sibling found by binary lifting
                // over the offsets.
                uint32_t lo = b;  // sibling of this lane's previous position (positions only grow)
                for (uint32_t base = 0; base < T; base += 64 * kFlatUnroll) {
                    uint32_t xs[kFlatUnroll], ws[kFlatUnroll];
#pragma unroll
                    for (int k = 0; k < kFlatUnroll; ++k) {
                        const uint32_t p = min(base + 64 * k + lane, T - 1);
#pragma unroll
                        for (uint32_t step = 32; step; step >>= 1) {
                            const uint32_t t = lo + step;
                            const uint32_t ot = fl.off[min(t, en - 1)];
                            lo = (t < en && ot <= p) ? t : lo;
                        }
                        xs[k] = a.col[fl.start[lo] + (p - fl.off[lo])];
                        ws[k] = fl.w[lo];
                    }
#pragma unroll
                    for (int k = 0; k < kFlatUnroll; ++k) {
                        const bool valid = base + 64 * k + lane < T;
                        const bool found = bsearch_global(P.p, P.n, xs[k]);
                        const uint32_t cls = pos_class_plus<S - 1>(xs[k], zs, ws[k]);
                        const uint32_t fm = 0u - uint32_t(found);
#pragma unroll
                        for (int i = 0; i < NS; ++i)
                            c[i] += valid && ((((cls_in[i] & fm) | (cls_out[i] & ~fm)) >> cls) & 1u);
                    }
                }
            }
#pragma unroll
            for (int i = 0; i < NS; ++i)
                if (dst[i] == -1) add_agg(lacc, int(agg[i]), neg[i] ? 0ull - c[i] : (unsigned long long)c[i]);
        };
        if (ns <= 1) run_stream(std::integral_constant<int, 1>());
        else if (ns <= 2) run_stream(std::integral_constant<int, 2>());
        else if (ns <= 4) run_stream(std::integral_constant<int, 4>());
        else run_stream(std::integral_constant<int, NC>());
    }
    // per-sibling parts: counts of an ancestor set, and the P part of P \ N(w)
    uint32_t extra[KZ_MAX_OPS + 1] = {0, 0, 0, 0, 0};  // SIZES: children, cached sizes
#pragma unroll
    for (int o = 0; o < KZ_MAX_OUTS; ++o) {
        if (o >= L.num_outs) break;
        if ((omode[o] & 3u) && !(omode[o] & 4u)) continue;
        const kz_out& out = L.outs[o];
        const SetView A = (omode[o] & 4u) ? P : resolve_shared(out.src);
        if (!sib) continue;
        const uint64_t cnt = count_classes<S>(A.p, A.n, zw, out.pred);
        if (cdest[o] == -1) add_agg(lacc, out.aggregator, cnt);
        else extra[0] += uint32_t(cnt);
    }
    if constexpr (SIZES) {
#pragma unroll
        for (int j = 0; j < KZ_MAX_OPS; ++j)
            if (omode[KZ_MAX_OUTS + j] & 4u) extra[1 + j] += P.n;
        wave_lds_sync();
        if (sib) {
            if (a.nchild) a.nchild[e_lane - a.begin] = fl.sib[0][lane] + extra[0];
#pragma unroll
            for (int j = 0; j < KZ_MAX_OPS; ++j)
                if (a.cached_mask & (1u << j))
                    a.slot_len_out[j][e_lane - a.begin] = j < L.num_ops ? fl.sib[1 + j][lane] + extra[1 + j] : 0u;
        }
    }
}

// FLAT: 0 = every embedding takes the general path; 1 = every list at size
// S only counts and takes the flat path; 2 = pass A of a size whose lists
// all take the flat path with per-sibling sizes (pass B takes the general
// path).  (One kernel holding both paths needs twice the registers of
// either.)
template <int S, int FLAT>
__global__ __launch_bounds__(256) void plan_level_kernel(LevelArgs a) {
    __shared__ uint32_t sHash[kWavesPerBlock][kHashCap];
    __shared__ FlatLds<FLAT == 2> sFlat[FLAT ? kWavesPerBlock : 1];
    static_assert(!FLAT || S >= 2, "the flat path needs an ancestor");
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t w = threadIdx.x >> 6;
    uint32_t* hk = sHash[w];
    const uint64_t gwave = uint64_t(blockIdx.x) * kWavesPerBlock + w;
    uint32_t* temp = a.temp + gwave * uint64_t(KZ_MAX_OPS) * a.temp_cap;
    HashCache hc{nullptr, 0};
    unsigned long long acc[KZ_MAX_AGGREGATORS];   // wave-uniform counts
    unsigned long long lacc[KZ_MAX_AGGREGATORS];  // per-lane counts (flat path)
#pragma unroll
    for (int i = 0; i < KZ_MAX_AGGREGATORS; ++i) acc[i] = lacc[i] = 0;
    while (true) {
        uint64_t first = 0;
        if (lane == 0) first = atomicAdd(reinterpret_cast<unsigned long long*>(a.cursor), (unsigned long long)kBatch);
        // (register allocation: readlane suits the flat form, shfl the general one)
        if constexpr (FLAT) first = readlane_u64(first, 0) + a.begin;
        else first = __shfl(first, 0) + a.begin;
        if (first >= a.end) break;
        const uint64_t last = first + kBatch < a.end ? first + kBatch : a.end;
        if constexpr (FLAT) {
            {
                // one embedding per lane; runs of lanes with the same parent and
                // step list are segments
                const uint64_t e = first + lane;
                const bool valid = e < last;
                uint32_t m[S];
                uint64_t idx[S + 1];
                uint32_t li = 0xFF;
                uint64_t parent = ~0ull;
                uint64_t wstart = 0;  // the new vertex's list
                uint32_t wlen = 0;
                if (valid) {
                    rebuild<S>(a, e, m, idx);
                    li = a.types[order_code(m, S)];
                    parent = idx[S - 1];
                    wstart = a.rowptr[m[S - 1]];
                    wlen = uint32_t(a.rowptr[m[S - 1] + 1] - wstart);
                } else {
#pragma unroll
                    for (int t = 0; t < S; ++t) m[t] = 0;
#pragma unroll
                    for (int t = 0; t <= S; ++t) idx[t] = 0;
                }
                const uint64_t pp = __shfl_up(parent, 1);
                const uint32_t pl = __shfl_up(li, 1);
                uint64_t heads = __ballot(valid && (lane == 0 || pp != parent || pl != li));
                const uint32_t nvalid = uint32_t(last - first);
                while (heads) {
                    const uint32_t b = uint32_t(__builtin_ctzll(heads));
                    heads &= heads - 1;
                    const uint32_t en = heads ? uint32_t(__builtin_ctzll(heads)) : nvalid;
                    const uint32_t sli = __builtin_amdgcn_readlane(li, b);
                    if (sli == 0xFF) {
                        if (lane == 0) atomicOr(a.err, 1u);
                        continue;
                    }
                    flat_segment<S, FLAT == 2>(a, a.lists[sli], b, en, m, idx, wstart, wlen, e, hk, hc, sFlat[w],
                                               acc, lacc);
                    // counting lists make no children and cache no prefix
                    if (FLAT == 1 && lane >= b && lane < en) {
                        if (a.nchild) a.nchild[e - a.begin] = 0;
#pragma unroll
                        for (int j = 0; j < KZ_MAX_OPS; ++j)
                            if (a.cached_mask & (1u << j)) a.slot_len_out[j][e - a.begin] = 0;
                    }
                }
                continue;
            }
        }
        if constexpr (!FLAT)
            for (uint64_t e = first; e < last; ++e) one_embedding(a, e, hk, hc, acc, temp);
    }
    if (a.mode == 0) {
#pragma unroll
        for (int i = 0; i < KZ_MAX_AGGREGATORS; ++i) {
            unsigned long long v = acc[i];
            if constexpr (FLAT) v += wave_sum_u64(lacc[i]);
            if (lane == 0 && v) atomicAdd(&a.agg[i], v);
        }
    }
}

// Smallest e in [0, n] with prefix[e] >= target (prefix has n+1 entries).
__global__ void chunk_bounds_kernel(const uint64_t* __restrict__ cost, uint64_t n, uint64_t budget, uint32_t nb,
                                    uint64_t* __restrict__ bounds) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k > nb) return;
    if (k == nb) {
        bounds[k] = n;
        return;
    }
    const uint64_t target = uint64_t(k) * budget;
    uint64_t lo = 0, hi = n;
    while (lo < hi) {
        const uint64_t mid = (lo + hi) >> 1;
        if (cost[mid] < target) lo = mid + 1;
        else hi = mid;
    }
    bounds[k] = lo;
}

// cost[e] = 8 * child_off[e] + 4 * sum_j slot_off[j][e]  (all exclusive prefixes, n+1 entries)
__global__ void chunk_cost_kernel(uint64_t n, const uint64_t* __restrict__ child_off, const uint64_t* __restrict__ s0,
                                  const uint64_t* __restrict__ s1, const uint64_t* __restrict__ s2,
                                  const uint64_t* __restrict__ s3, uint64_t* __restrict__ cost) {
    const uint64_t e = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e > n) return;
    uint64_t c = 8 * (child_off ? child_off[e] : 0);
    if (s0) c += 4 * s0[e];
    if (s1) c += 4 * s1[e];
    if (s2) c += 4 * s2[e];
    if (s3) c += 4 * s3[e];
    cost[e] = c;
}

struct CastU64 {
    __host__ __device__ uint64_t operator()(uint32_t x) const { return x; }
};

// Exclusive scan of u32 counts into u64 offsets with n+1 entries.
static int scan_u32(const uint32_t* in, uint64_t* out, uint64_t n, Scratch& tmp, hipStream_t s) {
    hipcub::TransformInputIterator<uint64_t, CastU64, const uint32_t*> it(in, CastU64());
    size_t bytes = 0;
    KZ_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, it, out, n + 1, s));
    if (int e = scratch_reserve(tmp, bytes)) return e;
    KZ_HIP(hipcub::DeviceScan::ExclusiveSum(tmp.ptr, bytes, it, out, n + 1, s));
    return KZ_OK;
}

// ----------------------------------------------------------------- host --

// Whether step list L at size s can take the flat path: every op is
// counted while streaming (not cached, read by no later op, feeding COUNT
// outputs only), intersects or subtracts one ancestor set P -- the same for
// every op -- and the new vertex's list N(m[s-1]); every output is a COUNT
// of an op's result, of N(m[s-1]) or of an ancestor set.
static bool flat_eligible(const kz_step_list& L, int s, uint32_t cached) {
    auto varying = [&](const kz_set_ref& r) { return r.kind == KZ_SET_NBR && int(r.a) == s - 1; };
    auto shared = [&](const kz_set_ref& r) {
        return (r.kind == KZ_SET_NBR && int(r.a) < s - 1) || (r.kind == KZ_SET_SLOT && int(r.a) < s);
    };
    auto same = [](const kz_set_ref& x, const kz_set_ref& y) { return x.kind == y.kind && x.a == y.a && x.b == y.b; };
    auto is_op = [&](const kz_set_ref& r, int j) { return r.kind == KZ_SET_SLOT && int(r.a) == s && int(r.b) == j; };
    kz_set_ref P{};
    for (int j = 0; j < L.num_ops; ++j) {
        const kz_op& op = L.ops[j];
        if (cached & (1u << j)) return false;
        for (int k = j + 1; k < L.num_ops; ++k)
            if (is_op(L.ops[k].lhs, j) || is_op(L.ops[k].rhs, j)) return false;
        bool read = false;
        for (int o = 0; o < L.num_outs; ++o)
            if (is_op(L.outs[o].src, j)) read = true;
        if (!read) return false;
        kz_set_ref sh;
        if (varying(op.lhs) && shared(op.rhs)) sh = op.rhs;
        else if (shared(op.lhs) && varying(op.rhs)) sh = op.lhs;
        else return false;
        if (op.op != KZ_OP_INTERSECT && op.op != KZ_OP_SUBTRACT) return false;
        if (j == 0) P = sh;
        else if (!same(P, sh)) return false;
    }
    for (int o = 0; o < L.num_outs; ++o) {
        const kz_out& out = L.outs[o];
        if (out.kind != KZ_OUT_COUNT) return false;
        const kz_set_ref& r = out.src;
        const bool op_result = r.kind == KZ_SET_SLOT && int(r.a) == s && int(r.b) < L.num_ops;
        if (!op_result && !varying(r) && !shared(r)) return false;
    }
    return true;
}

// Whether step list L at size s can take the flat path with per-sibling
// sizes (pass A of a size that also scatters or caches): every op
// intersects or subtracts one ancestor set P -- the same for all ops -- and
// N(m[s-1]); every output is a COUNT or a SCATTER of an op's result, of
// N(m[s-1]) or of an ancestor set.
static bool flat_sizes_eligible(const kz_step_list& L, int s) {
    auto varying = [&](const kz_set_ref& r) { return r.kind == KZ_SET_NBR && int(r.a) == s - 1; };
    auto shared = [&](const kz_set_ref& r) {
        return (r.kind == KZ_SET_NBR && int(r.a) < s - 1) || (r.kind == KZ_SET_SLOT && int(r.a) < s);
    };
    auto same = [](const kz_set_ref& x, const kz_set_ref& y) { return x.kind == y.kind && x.a == y.a && x.b == y.b; };
    kz_set_ref P{};
    for (int j = 0; j < L.num_ops; ++j) {
        const kz_op& op = L.ops[j];
        if (op.op != KZ_OP_INTERSECT && op.op != KZ_OP_SUBTRACT) return false;
        kz_set_ref sh;
        if (varying(op.lhs) && shared(op.rhs)) sh = op.rhs;
        else if (shared(op.lhs) && varying(op.rhs)) sh = op.lhs;
        else return false;
        if (j == 0) P = sh;
        else if (!same(P, sh)) return false;
    }
    for (int o = 0; o < L.num_outs; ++o) {
        const kz_out& out = L.outs[o];
        if (out.kind != KZ_OUT_COUNT && out.kind != KZ_OUT_SCATTER) return false;
        const kz_set_ref& r = out.src;
        const bool op_result = r.kind == KZ_SET_SLOT && int(r.a) == s && int(r.b) < L.num_ops;
        if (!op_result && !varying(r) && !shared(r)) return false;
    }
    return true;
}

struct LevelScratch {
    Scratch nchild, child_off, slot_len[KZ_MAX_OPS], slot_off[KZ_MAX_OPS], cost, bounds;
    Scratch child_parent, child_vertex, slot_data[KZ_MAX_OPS];
};

struct Executor {
    DeviceGraph& g;
    const PlanDesc& p;
    kz_step_list* d_lists = nullptr;
    uint8_t* d_fuse = nullptr;
    uint8_t* d_types = nullptr;
    uint32_t* d_cursor = nullptr;
    unsigned long long* d_agg = nullptr;
    uint32_t* d_err = nullptr;
    Scratch temp, scan_tmp;
    LevelScratch ls[KZ_MAX_PATTERN + 1];
    uint32_t cached_mask[KZ_MAX_PATTERN + 1] = {};
    bool has_scatter[KZ_MAX_PATTERN + 1] = {};
    int flat_level[KZ_MAX_PATTERN + 1] = {};  // every list at this size takes the flat path: 1 counting, 2 with sizes
    uint32_t grid = 0;
    uint32_t temp_cap = 0;
    kz_run_stats* st = nullptr;
    uint64_t peak = 0;
    uint64_t in_use = 0;

    Executor(DeviceGraph& g_, const PlanDesc& p_) : g(g_), p(p_) {}
    ~Executor() {
        hipFree(d_lists);
        hipFree(d_fuse);
        hipFree(d_types);
        hipFree(d_cursor);
        hipFree(d_agg);
        hipFree(d_err);
        auto rel = [](Scratch& s) {
            if (s.ptr) hipFree(s.ptr);
            s.ptr = nullptr;
        };
        rel(temp);
        rel(scan_tmp);
        for (auto& l : ls) {
            rel(l.nchild);
            rel(l.child_off);
            rel(l.cost);
            rel(l.bounds);
            rel(l.child_parent);
            rel(l.child_vertex);
            for (int j = 0; j < KZ_MAX_OPS; ++j) {
                rel(l.slot_len[j]);
                rel(l.slot_off[j]);
                rel(l.slot_data[j]);
            }
        }
    }

    int setup() {
        const int K = int(p.pattern_size);
        std::vector<int> list_at(std::max<uint32_t>(1, p.num_step_lists), -1);
        // which slots of size s are read by larger sizes (must be cached)
        for (uint32_t li = 0; li < p.num_step_lists; ++li) {
            const kz_step_list& L = p.lists[li];
            auto mark = [&](const kz_set_ref& r, int at) {
                if (r.kind == KZ_SET_SLOT && r.a < at) cached_mask[r.a] |= 1u << r.b;
            };
            int at = -1;
            for (int s = 1; s < K && at < 0; ++s)
                for (int c = 0; c < 256; ++c)
                    if (p.type_table[s][c] == li) {
                        at = s;
                        break;
                    }
            if (at < 0) continue;
            for (int j = 0; j < L.num_ops; ++j) {
                mark(L.ops[j].lhs, at);
                mark(L.ops[j].rhs, at);
            }
            for (int o = 0; o < L.num_outs; ++o) {
                mark(L.outs[o].src, at);
                if (L.outs[o].kind == KZ_OUT_SCATTER) has_scatter[at] = true;
            }
            list_at[li] = at;
        }
        // ops whose result feeds only COUNT outputs of the same step list
        // (not cached, not read by a later op) are counted while streaming
        std::vector<uint8_t> fuse(std::max<uint32_t>(1, p.num_step_lists), 0);
        for (uint32_t li = 0; li < p.num_step_lists; ++li) {
            const int at = list_at[li];
            if (at < 0) continue;
            const kz_step_list& L = p.lists[li];
            for (int j = 0; j < L.num_ops; ++j) {
                if (cached_mask[at] & (1u << j)) continue;
                auto is_j = [&](const kz_set_ref& r) { return r.kind == KZ_SET_SLOT && r.a == at && r.b == j; };
                bool ok = true, any = false;
                for (int k = j + 1; k < L.num_ops; ++k) ok &= !is_j(L.ops[k].lhs) && !is_j(L.ops[k].rhs);
                for (int o = 0; o < L.num_outs; ++o)
                    if (is_j(L.outs[o].src)) {
                        any = true;
                        ok &= L.outs[o].kind == KZ_OUT_COUNT;
                    }
                if (ok && any) fuse[li] |= uint8_t(1u << j);
            }
        }
        KZ_HIP(hipMalloc(&d_fuse, fuse.size()));
        KZ_HIP(hipMemcpy(d_fuse, fuse.data(), fuse.size(), hipMemcpyHostToDevice));
        // KZ_PLAN_FLAT=0 keeps every size on the general path (for comparisons);
        // stream offsets are 32-bit: 64 lists of fewer than 2^26 entries
        const char* env = std::getenv("KZ_PLAN_FLAT");
        const bool flat_on = !(env && env[0] == '0') && g.max_degree < (1u << 26);
        for (int s = 2; s < K && flat_on; ++s) {
            bool all = true, all_sizes = true;
            for (int c = 0; c < 256; ++c) {
                const uint8_t li = p.type_table[s][c];
                if (li == 0xFF) continue;
                all &= flat_eligible(p.lists[li], s, cached_mask[s]);
                all_sizes &= flat_sizes_eligible(p.lists[li], s);
            }
            flat_level[s] = all ? 1 : (all_sizes ? 2 : 0);
        }
        KZ_HIP(hipMalloc(&d_lists, sizeof(kz_step_list) * std::max<uint32_t>(1, p.num_step_lists)));
        KZ_HIP(hipMalloc(&d_types, sizeof(p.type_table)));
        KZ_HIP(hipMalloc(&d_cursor, 8));
        KZ_HIP(hipMalloc(&d_agg, sizeof(unsigned long long) * KZ_MAX_AGGREGATORS));
        KZ_HIP(hipMalloc(&d_err, 4));
        hipStream_t s = rt().stream;
        KZ_HIP(hipMemcpyAsync(d_lists, p.lists, sizeof(kz_step_list) * p.num_step_lists, hipMemcpyHostToDevice, s));
        KZ_HIP(hipMemcpyAsync(d_types, p.type_table, sizeof(p.type_table), hipMemcpyHostToDevice, s));
        KZ_HIP(hipMemsetAsync(d_agg, 0, sizeof(unsigned long long) * KZ_MAX_AGGREGATORS, s));
        KZ_HIP(hipMemsetAsync(d_err, 0, 4, s));
        grid = uint32_t(rt().num_cus) * 4;
        temp_cap = std::max<uint32_t>(g.max_degree, 1);
        if (int e = scratch_reserve(temp, size_t(grid) * kWavesPerBlock * KZ_MAX_OPS * temp_cap * 4)) return e;
        return KZ_OK;
    }

    int launch(LevelArgs& la) {
        hipStream_t s = rt().stream;
        KZ_HIP(hipMemsetAsync(d_cursor, 0, 8, s));
        la.cursor = d_cursor;
        la.agg = d_agg;
        la.err = d_err;
        la.lists = d_lists;
        la.fuse = d_fuse;
        la.types = d_types + la.s * 256;
        la.rowptr = g.rowptr;
        la.col = g.col;
        la.temp = static_cast<uint32_t*>(temp.ptr);
        la.temp_cap = temp_cap;
        // a counting flat size has nothing to write: its pass A is the whole
        // level; a flat size with sizes writes in pass B on the general path
        int f = flat_level[la.s];
        if (f == 1 && la.mode != 0) return KZ_OK;
        if (la.mode != 0) f = 0;
        const dim3 gd(grid), bd(64 * kWavesPerBlock);
#define KZ_LEVEL_CASE(S_)                                                                            \
    case S_:                                                                                         \
        if (f == 1) hipLaunchKernelGGL((plan_level_kernel<S_, 1>), gd, bd, 0, s, la);                \
        else if (f == 2) hipLaunchKernelGGL((plan_level_kernel<S_, 2>), gd, bd, 0, s, la);           \
        else hipLaunchKernelGGL((plan_level_kernel<S_, 0>), gd, bd, 0, s, la);                       \
        break;
        switch (la.s) {
            case 1: hipLaunchKernelGGL((plan_level_kernel<1, 0>), gd, bd, 0, s, la); break;
            KZ_LEVEL_CASE(2)
            KZ_LEVEL_CASE(3)
            KZ_LEVEL_CASE(4)
            default: KZ_FAIL(KZ_ERR_INVALID, "embedding size %d outside 1..%d", la.s, KZ_MAX_PATTERN - 1);
        }
#undef KZ_LEVEL_CASE
        KZ_HIP(hipGetLastError());
        if (st) st->launches++;
        return KZ_OK;
    }

    // Extend every embedding of size s held in lv[s] (all ancestors in lv).
    int run_level(int s, DevLevel* lv) {
        const int K = int(p.pattern_size);
        const uint64_t n = lv[s].n;
        if (!n) return KZ_OK;
        if (st) st->embeddings[s] += n;
        hipStream_t stream = rt().stream;
        LevelScratch& L = ls[s];
        const bool deeper = (s + 1 < K) && has_scatter[s];
        LevelArgs la;
        std::memset(&la, 0, sizeof(la));
        for (int t = 1; t <= s; ++t) la.lv[t] = lv[t];
        la.s = s;
        la.begin = 0;
        la.end = n;
        la.mode = 0;
        const uint32_t cm = deeper ? cached_mask[s] : 0u;
        la.cached_mask = cm;
        if (deeper) {
            // n + 1 entries: the scans read one element past the last embedding
            if (int e = scratch_reserve(L.nchild, (n + 1) * 4)) return e;
            la.nchild = static_cast<uint32_t*>(L.nchild.ptr);
            for (int j = 0; j < KZ_MAX_OPS; ++j)
                if (cm & (1u << j)) {
                    if (int e = scratch_reserve(L.slot_len[j], (n + 1) * 4)) return e;
                    la.slot_len_out[j] = static_cast<uint32_t*>(L.slot_len[j].ptr);
                }
        }
        if (int e = launch(la)) return e;
        if (!deeper) return KZ_OK;
        // scans
        if (int e = scratch_reserve(L.child_off, (n + 1) * 8)) return e;
        uint64_t* child_off = static_cast<uint64_t*>(L.child_off.ptr);
        if (int e = scan_u32(la.nchild, child_off, n, scan_tmp, stream)) return e;
        uint64_t* slot_off[KZ_MAX_OPS] = {};
        for (int j = 0; j < KZ_MAX_OPS; ++j)
            if (cm & (1u << j)) {
                if (int e = scratch_reserve(L.slot_off[j], (n + 1) * 8)) return e;
                slot_off[j] = static_cast<uint64_t*>(L.slot_off[j].ptr);
                if (int e = scan_u32(la.slot_len_out[j], slot_off[j], n, scan_tmp, stream)) return e;
            }
        if (int e = scratch_reserve(L.cost, (n + 1) * 8)) return e;
        uint64_t* cost = static_cast<uint64_t*>(L.cost.ptr);
        hipLaunchKernelGGL(chunk_cost_kernel, dim3((n + 1 + 255) / 256), dim3(256), 0, stream, n, child_off,
                           slot_off[0], slot_off[1], slot_off[2], slot_off[3], cost);
        uint64_t total = 0;
        KZ_HIP(hipMemcpyAsync(&total, cost + n, 8, hipMemcpyDeviceToHost, stream));
        KZ_HIP(hipStreamSynchronize(stream));
        const uint64_t budget = std::max<uint64_t>(g_level_budget, 1 << 20);
        uint32_t nb = uint32_t(std::max<uint64_t>(1, (total + budget - 1) / budget));
        if (int e = scratch_reserve(L.bounds, (size_t(nb) + 1) * 8)) return e;
        uint64_t* d_bounds = static_cast<uint64_t*>(L.bounds.ptr);
        hipLaunchKernelGGL(chunk_bounds_kernel, dim3((nb + 1 + 255) / 256), dim3(256), 0, stream, cost, n, budget, nb,
                           d_bounds);
        std::vector<uint64_t> bounds(size_t(nb) + 1);
        KZ_HIP(hipMemcpyAsync(bounds.data(), d_bounds, (size_t(nb) + 1) * 8, hipMemcpyDeviceToHost, stream));
        KZ_HIP(hipStreamSynchronize(stream));
        bounds[0] = 0;
        bounds[nb] = n;
        for (uint32_t c = 0; c < nb; ++c) {
            const uint64_t p0 = bounds[c], p1 = bounds[c + 1];
            if (p1 <= p0) continue;
            // sizes of this chunk
            uint64_t h[2 + 2 * KZ_MAX_OPS];
            KZ_HIP(hipMemcpyAsync(&h[0], child_off + p0, 8, hipMemcpyDeviceToHost, stream));
            KZ_HIP(hipMemcpyAsync(&h[1], child_off + p1, 8, hipMemcpyDeviceToHost, stream));
            for (int j = 0; j < KZ_MAX_OPS; ++j)
                if (slot_off[j]) {
                    KZ_HIP(hipMemcpyAsync(&h[2 + 2 * j], slot_off[j] + p0, 8, hipMemcpyDeviceToHost, stream));
                    KZ_HIP(hipMemcpyAsync(&h[3 + 2 * j], slot_off[j] + p1, 8, hipMemcpyDeviceToHost, stream));
                }
            KZ_HIP(hipStreamSynchronize(stream));
            const uint64_t nchild = h[1] - h[0];
            if (st) st->chunks++;
            if (int e = scratch_reserve(L.child_parent, std::max<uint64_t>(1, nchild) * 4)) return e;
            if (int e = scratch_reserve(L.child_vertex, std::max<uint64_t>(1, nchild) * 4)) return e;
            LevelArgs lb = la;
            lb.mode = 1;
            lb.begin = p0;
            lb.end = p1;
            lb.nchild = nullptr;
            lb.child_off = child_off + p0;
            lb.child_base = h[0];
            lb.child_parent = static_cast<uint32_t*>(L.child_parent.ptr);
            lb.child_vertex = static_cast<uint32_t*>(L.child_vertex.ptr);
            DevLevel cur = lv[s];
            for (int j = 0; j < KZ_MAX_OPS; ++j) {
                if (!slot_off[j]) continue;
                const uint64_t len = h[3 + 2 * j] - h[2 + 2 * j];
                if (int e = scratch_reserve(L.slot_data[j], std::max<uint64_t>(1, len) * 4)) return e;
                lb.slot_off_out[j] = slot_off[j] + p0;
                lb.slot_base_out[j] = h[2 + 2 * j];
                lb.slot_data_out[j] = static_cast<uint32_t*>(L.slot_data[j].ptr);
                cur.slot_off[j] = slot_off[j];
                cur.slot_len[j] = la.slot_len_out[j];
                cur.slot_data[j] = lb.slot_data_out[j];
                cur.slot_base[j] = h[2 + 2 * j];
            }
            if (int e = launch(lb)) return e;
            track();
            DevLevel next[KZ_MAX_PATTERN + 1];
            for (int t = 1; t <= s; ++t) next[t] = lv[t];
            next[s] = cur;
            next[s + 1] = DevLevel();
            next[s + 1].n = nchild;
            next[s + 1].parent = lb.child_parent;
            next[s + 1].vertex = lb.child_vertex;
            if (int e = run_level(s + 1, next)) return e;
        }
        return KZ_OK;
    }

    void track() {
        uint64_t tot = temp.size + scan_tmp.size;
        for (auto& l : ls) {
            tot += l.nchild.size + l.child_off.size + l.cost.size + l.bounds.size + l.child_parent.size +
                   l.child_vertex.size;
            for (int j = 0; j < KZ_MAX_OPS; ++j) tot += l.slot_len[j].size + l.slot_off[j].size + l.slot_data[j].size;
        }
        peak = std::max(peak, tot);
    }
};

int plan_execute(DeviceGraph& g, const PlanDesc& p, bool, uint64_t* raw_counts, kz_run_stats* st) {
    if (int e = plan_validate(p)) return e;
    if (g.n == 0) {
        for (uint32_t i = 0; i < p.num_aggregators; ++i) raw_counts[i] = 0;
        return KZ_OK;
    }
    Executor x(g, p);
    x.st = st;
    if (int e = x.setup()) return e;
    hipStream_t s = rt().stream;
    hipEvent_t ev0, ev1;
    KZ_HIP(hipEventCreate(&ev0));
    KZ_HIP(hipEventCreate(&ev1));
    KZ_HIP(hipEventRecord(ev0, s));
    DevLevel lv[KZ_MAX_PATTERN + 1];
    lv[1].n = g.n;
    lv[1].root_base = 0;
    int e = x.run_level(1, lv);
    KZ_HIP(hipEventRecord(ev1, s));
    unsigned long long agg[KZ_MAX_AGGREGATORS];
    uint32_t err = 0;
    KZ_HIP(hipMemcpyAsync(agg, x.d_agg, sizeof(agg), hipMemcpyDeviceToHost, s));
    KZ_HIP(hipMemcpyAsync(&err, x.d_err, 4, hipMemcpyDeviceToHost, s));
    KZ_HIP(hipStreamSynchronize(s));
    float ms = 0;
    KZ_HIP(hipEventElapsedTime(&ms, ev0, ev1));
    hipEventDestroy(ev0);
    hipEventDestroy(ev1);
    if (e) return e;
    if (err) KZ_FAIL(KZ_ERR_UNSUPPORTED, "an embedding reached an order type the plan does not define");
    for (uint32_t i = 0; i < p.num_aggregators && i < KZ_MAX_AGGREGATORS; ++i) raw_counts[i] = agg[i];
    x.track();
    if (st) {
        st->device_ms = ms;
        st->peak_bytes = x.peak;
    }
    return KZ_OK;
}

// ------------------------------------------------------------ plan utils --

static uint32_t ipow(uint32_t b, int e) {
    uint32_t r = 1;
    while (e-- > 0) r *= b;
    return r;
}

int plan_validate(const PlanDesc& p) {
    const int K = int(p.pattern_size);
    if (K < 2 || K > KZ_MAX_PATTERN) KZ_FAIL(KZ_ERR_INVALID, "pattern_size %d outside 2..%d", K, KZ_MAX_PATTERN);
    if (p.num_aggregators > KZ_MAX_AGGREGATORS || p.num_step_lists > KZ_MAX_STEP_LISTS)
        KZ_FAIL(KZ_ERR_INVALID, "too many aggregators or step lists");
    for (int s = 1; s < K; ++s) {
        const uint32_t ncode = ipow(uint32_t(s), s);
        for (uint32_t c = 0; c < 256; ++c) {
            const uint8_t li = p.type_table[s][c];
            if (li == 0xFF) continue;
            if (c >= ncode) KZ_FAIL(KZ_ERR_INVALID, "order code %u invalid for size %d", c, s);
            if (li >= p.num_step_lists) KZ_FAIL(KZ_ERR_INVALID, "step list %u out of range", li);
            const kz_step_list& L = p.lists[li];
            if (L.num_ops > KZ_MAX_OPS || L.num_outs > KZ_MAX_OUTS) KZ_FAIL(KZ_ERR_INVALID, "step list too long");
            auto ok_ref = [&](const kz_set_ref& r, int j_limit) {
                if (r.kind == KZ_SET_NBR) return int(r.a) < s;
                if (r.kind == KZ_SET_SLOT) {
                    if (r.a < 1 || int(r.a) > s || r.b >= KZ_MAX_OPS) return false;
                    if (int(r.a) == s) return int(r.b) < j_limit;
                    return true;
                }
                return false;
            };
            for (int j = 0; j < L.num_ops; ++j) {
                const kz_op& o = L.ops[j];
                if (o.op != KZ_OP_INTERSECT && o.op != KZ_OP_SUBTRACT) KZ_FAIL(KZ_ERR_INVALID, "bad op");
                if (!ok_ref(o.lhs, j) || !ok_ref(o.rhs, j)) KZ_FAIL(KZ_ERR_INVALID, "bad operand at size %d op %d", s, j);
            }
            for (int o = 0; o < L.num_outs; ++o) {
                const kz_out& out = L.outs[o];
                if (!ok_ref(out.src, L.num_ops)) KZ_FAIL(KZ_ERR_INVALID, "bad output source at size %d", s);
                if (out.kind == KZ_OUT_COUNT && out.aggregator >= p.num_aggregators)
                    KZ_FAIL(KZ_ERR_INVALID, "aggregator %u out of range", out.aggregator);
                if (out.kind == KZ_OUT_SCATTER && s + 1 >= K)
                    KZ_FAIL(KZ_ERR_INVALID, "scatter at the last size %d", s);
                if (out.kind != KZ_OUT_COUNT && out.kind != KZ_OUT_SCATTER) KZ_FAIL(KZ_ERR_INVALID, "bad output kind");
            }
        }
    }
    return KZ_OK;
}

// kautomine k-clique schedule: size 1 scatters v > v0 from N(v0); size 2
// caches N(v0) & N(v1) (four_clique.cc:88-115); every deeper size
// intersects the parent's cached prefix with N(v_{s-1}) and scatters or
// counts v > v_{s-1} (four_clique.cc:116-142, five_clique.cc:116-175).
int plan_clique(int k, PlanDesc* out) {
    if (k < 3 || k > KZ_MAX_PATTERN) KZ_FAIL(KZ_ERR_UNSUPPORTED, "clique size %d", k);
    std::memset(out, 0, sizeof(*out));
    std::memset(out->type_table, 0xFF, sizeof(out->type_table));
    out->pattern_size = uint32_t(k);
    out->num_aggregators = 1;
    out->multiplicity[0] = 1;
    for (int s = 1; s < k; ++s) {
        kz_step_list& L = out->lists[s - 1];
        if (s >= 2) {
            L.num_ops = 1;
            L.ops[0].op = KZ_OP_INTERSECT;
            if (s == 2) {
                L.ops[0].lhs = kz_set_ref{KZ_SET_NBR, 0, 0, 0};
            } else {
                L.ops[0].lhs = kz_set_ref{KZ_SET_SLOT, uint8_t(s - 1), 0, 0};
            }
            L.ops[0].rhs = kz_set_ref{KZ_SET_NBR, uint8_t(s - 1), 0, 0};
        }
        L.num_outs = 1;
        kz_out& o = L.outs[0];
        o.kind = (s + 1 == k) ? KZ_OUT_COUNT : KZ_OUT_SCATTER;
        o.aggregator = 0;
        o.child_active = 1;
        o.src = s == 1 ? kz_set_ref{KZ_SET_NBR, 0, 0, 0} : kz_set_ref{KZ_SET_SLOT, uint8_t(s), 0, 0};
        o.pred = 1u << (2 * s);  // v above every matched vertex
        // increasing order type: dense ranks 0, 1, ..., s-1
        uint32_t code = 0, pw = 1;
        for (int j = 0; j < s; ++j) {
            code += uint32_t(j) * pw;
            pw *= uint32_t(s);
        }
        out->type_table[s][code] = uint8_t(s - 1);
    }
    out->num_step_lists = uint32_t(k - 1);
    return KZ_OK;
}

}  // namespace kz

using namespace kz;

extern "C" {

int kz_plan_clique(int k, kz_plan_desc* out) {
    if (!out) KZ_FAIL(KZ_ERR_INVALID, "null argument");
    return plan_clique(k, out);
}

int kz_plan_validate(const kz_plan_desc* plan) {
    if (!plan) KZ_FAIL(KZ_ERR_INVALID, "null argument");
    return plan_validate(*plan);
}

int kz_count_plan(kz_graph g, const kz_plan_desc* plan, uint64_t* raw_counts, uint32_t num_counts,
                  kz_run_stats* stats) {
    if (!g || !plan || !raw_counts) KZ_FAIL(KZ_ERR_INVALID, "null argument");
    if (num_counts < plan->num_aggregators) KZ_FAIL(KZ_ERR_INVALID, "raw_counts holds fewer than num_aggregators");
    if (int e = ensure_ready()) return e;
    auto t0 = std::chrono::steady_clock::now();
    kz_run_stats st;
    std::memset(&st, 0, sizeof(st));
    uint64_t tmp[KZ_MAX_AGGREGATORS] = {0};
    if (int e = plan_execute(g->g, *plan, false, tmp, &st)) return e;
    for (uint32_t i = 0; i < plan->num_aggregators; ++i) raw_counts[i] = tmp[i];
    st.wall_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    if (stats) *stats = st;
    return KZ_OK;
}

int kz_set_level_budget(uint64_t bytes) {
    if (bytes < (uint64_t(1) << 20)) KZ_FAIL(KZ_ERR_INVALID, "level budget below 1 MiB");
    g_level_budget = bytes;
    return KZ_OK;
}

}  // extern "C"
