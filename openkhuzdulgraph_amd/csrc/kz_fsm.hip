// Labelled frequent subgraph mining: 3- and 4-vertex patterns with MNI
// (minimum image) support, on one MI355X.
//
// Reference path replaced: kautomine/three_fsm.cc mining_frequent_patterns
// (:688-695): frequent labels (setup, :255-275), then frequent edges, wedges,
// triangles, four-chains and three-stars, each level's candidates built from
// the frequent patterns of the levels before (CandidateLabeledPatternGenerator,
// :276-604) and each candidate run through the engine with roots restricted
// to its first label (engine.cc:1459-1465), its apply() (:52-250) inserting
// the matched vertices into a DistributedPatternDomain (engine.h:405-580);
// the support is the smallest of the per-position domains.
//
// MI355X design (no counterpart in the reference):
//   * a domain depends only on which vertices take part in SOME embedding at
//     each position, not on how embeddings are enumerated, so each template
//     is decided by existence rules over label-filtered neighbourhoods: the
//     neighbour lists are stored label-sorted (ids ascending within a label),
//     so N_l(v), the neighbours of v with label l, is one contiguous segment
//     whose bounds are two words of a per-vertex table of label offsets
//     (L + 1 u32 per vertex, written by the same sort);
//   * one wave per root vertex (the vertices of the root label); domains are
//     bitmaps in HBM (one per candidate and position, set with atomicOr) and
//     their sizes come from one popcount pass per level;
//   * a level's candidates are all known before it runs, so every candidate
//     of a level is launched back to back into its own bitmaps and the level
//     synchronises once.
// Existence rules (labels l0..l3 of pattern positions 0..3; c_l(v) = |N_l(v)|):
//   edge      (v0-v1): D0 = {v0 : N_l1(v0) non-empty}, D1 = the union of those
//             N_l1(v0);
//   wedge     (v1-v0-v2): v0 qualifies iff c_l1 >= 1 and c_l2 >= 1 + [l1 = l2];
//             then v0 in D0 and all of N_l1(v0), N_l2(v0) in D1, D2 (any of
//             them has a partner of the other label distinct from itself);
//   triangle  per edge (v0, v1): T = N_l2(v0) & N_l2(v1); if non-empty, v0,
//             v1 and all of T join D0, D1, D2;
//   four-chain (v2-v0-v1-v3) per edge (v0, v1): P2 = N_l2(v0) \ {v1},
//             P3 = N_l3(v1) \ {v0}; a pair v2 != v3 exists iff both are
//             non-empty and not both the same single vertex.  Then v0, v1 join
//             D0, D1; v2 in P2 joins D2 unless P3 = {v2}, so D2 gains
//             N_l2(v0) minus the intersection over qualifying v1 of the
//             excluded set {v1} + ({e3} if |P3| = 1) -- at most 2 ids; D3
//             likewise from the v1 side;
//   three-star (center v1, leaves v0, v2, v3 with labels l0, l2, l3): v1
//             qualifies iff c_l(v1) >= the multiplicity of l among (l0, l2,
//             l3) for each of them; then v1 joins D1 and all of N_l0(v1),
//             N_l2(v1), N_l3(v1) join D0, D2, D3.

#include <algorithm>
#include <chrono>
#include <cstring>
#include <set>
#include <vector>

#include "kz_device.h"
#include "kz_kernels.h"

namespace kz {

constexpr uint32_t kNone = 0xFFFFFFFFu;

// Device state of a labelled graph (kz_graph_set_labels).
struct FsmState {
    uint32_t n = 0, L = 0;
    uint32_t* lab = nullptr;    // labels[n]
    uint32_t* lcol = nullptr;   // neighbour lists sorted by (label, id), same rowptr
    uint32_t* loff = nullptr;   // per vertex, L + 1 offsets of the label segments in its row
    uint32_t* vlist = nullptr;  // vertices grouped by label
    std::vector<uint32_t> voff;    // host: vlist offsets per label (L + 1)
    uint64_t bytes = 0;
    ~FsmState() {
        hipFree(lab);
        hipFree(lcol);
        hipFree(loff);
        hipFree(vlist);
    }
};

namespace {

// One wave per row: stable counting sort of the row's entries by label
// (labels < 64), through a per-wave LDS histogram and, chunk by chunk, one
// ballot per distinct label in the chunk; the histogram's exclusive prefix is
// the row's label offsets (loff[v * (L + 1) + l], l = 0..L).
__global__ __launch_bounds__(256) void label_sort_kernel(const uint64_t* __restrict__ rowptr,
                                                         const uint32_t* __restrict__ col,
                                                         const uint32_t* __restrict__ lab, uint32_t n, uint32_t L,
                                                         uint32_t* __restrict__ lcol, uint32_t* __restrict__ loff) {
    __shared__ uint32_t sBase[4][64];
    const uint32_t lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    uint32_t* base = sBase[w];
    const uint64_t lt = lane ? (~0ull >> (64 - lane)) : 0ull;
    for (uint32_t v = blockIdx.x * 4 + w; v < n; v += gridDim.x * 4) {
        const uint64_t b = rowptr[v], e = rowptr[v + 1];
        base[lane] = 0;
        wave_lds_sync();
        for (uint64_t k = b + lane; k < e; k += 64) atomicAdd(&base[lab[col[k]]], 1u);
        wave_lds_sync();
        const uint32_t c = base[lane];
        uint32_t excl;
        {
            uint32_t x = c;
            for (int o = 1; o < 64; o <<= 1) {
                const uint32_t y = __shfl_up(x, o);
                if (lane >= uint32_t(o)) x += y;
            }
            excl = x - c;
        }
        wave_lds_sync();
        base[lane] = excl;
        if (lane <= L) loff[size_t(v) * (L + 1) + lane] = excl;
        wave_lds_sync();
        for (uint64_t k0 = b; k0 < e; k0 += 64) {
            const uint64_t k = k0 + lane;
            const bool live = k < e;
            const uint32_t x = live ? col[k] : 0u;
            const uint32_t l = live ? lab[x] : 64u;
            unsigned long long todo = __ballot(live);
            while (todo) {
                const uint32_t lead = uint32_t(__builtin_ctzll(todo));
                const uint32_t ll = wave_lane(l, lead);
                const unsigned long long m = __ballot(l == ll);
                const uint32_t at = base[ll];
                if (l == ll) {
                    const uint64_t pos = b + at + uint32_t(__popcll(m & lt));
                    lcol[pos] = x;
                }
                wave_lds_sync();
                if (lane == lead) base[ll] = at + uint32_t(__popcll(m));
                wave_lds_sync();
                todo &= ~m;
            }
        }
        wave_lds_sync();
    }
}

struct Seg {
    uint64_t b, e;
};

// N_l(v): the entries of label l in v's label-sorted list.
__device__ __forceinline__ Seg seg_of(const uint64_t* __restrict__ rowptr, const uint32_t* __restrict__ loff,
                                      uint32_t L1, uint32_t v, uint32_t l) {
    const uint64_t b = ldg(rowptr, v);
    const uint32_t* o = loff + size_t(v) * L1 + l;
    return Seg{b + ldg(o, 0), b + ldg(o, 1)};
}

__device__ __forceinline__ void mark(uint32_t* bm, uint32_t v) { atomicOr(&bm[v >> 5], 1u << (v & 31)); }

// Marks every id of lcol[s.b, s.e) except x1, x2 (kNone: none) across the wave.
__device__ __forceinline__ void mark_seg(uint32_t* bm, const uint32_t* __restrict__ lcol, Seg s, uint32_t x1,
                                         uint32_t x2) {
    for (uint64_t k = s.b + (threadIdx.x & 63); k < s.e; k += 64) {
        const uint32_t w = ldg(lcol, k);
        if (w != x1 && w != x2) mark(bm, w);
    }
}

// The single id of lcol[s.b, s.e) other than `skip` (the segment holds one
// such id, plus `skip` or not).
__device__ __forceinline__ uint32_t only_other(const uint32_t* __restrict__ lcol, Seg s, uint32_t skip) {
    const uint32_t a = ldg(lcol, s.b);
    return (a != skip || s.e - s.b == 1) ? a : ldg(lcol, s.b + 1);
}

// Position of x in the sorted lcol[s.b, s.e), or kNone.
__device__ __forceinline__ bool seg_has(const uint32_t* __restrict__ lcol, Seg s, uint32_t x) {
    uint64_t a = s.b, z = s.e;
    while (a < z) {
        const uint64_t m = (a + z) >> 1;
        if (ldg(lcol, m) < x) a = m + 1;
        else z = m;
    }
    return a < s.e && ldg(lcol, a) == x;
}

struct FsmArgs {
    const uint64_t* rowptr;
    const uint32_t* lcol;
    const uint32_t* loff;
    uint32_t L1;  // labels + 1: the offset table's row length
    const uint32_t* roots;
    uint32_t nroots;
    uint32_t l[4];
    uint32_t* bm;  // 4 bitmaps of W words
    uint32_t W;
};

// Running intersection of wave lanes' exclusion sets {x1, x2} over lanes
// with ok set (i1, i2 uniform; `any` set once a first set was taken).
__device__ __forceinline__ void fold_exclusions(bool ok, uint32_t x1, uint32_t x2, bool& any, uint32_t& i1,
                                                uint32_t& i2) {
    const unsigned long long okm = __ballot(ok);
    if (!okm) return;
    if (!any) {
        const int f = __builtin_ctzll(okm);
        i1 = wave_lane(x1, uint32_t(f));
        i2 = wave_lane(x2, uint32_t(f));
        any = true;
    }
    if (i1 != kNone && __ballot(ok && x1 != i1 && x2 != i1)) i1 = kNone;
    if (i2 != kNone && __ballot(ok && x1 != i2 && x2 != i2)) i2 = kNone;
}

template <int T>
__global__ __launch_bounds__(256) void fsm_domain_kernel(FsmArgs a) {
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t wid = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    const uint32_t nw = gridDim.x * (blockDim.x >> 6);
    uint32_t* D0 = a.bm;
    uint32_t* D1 = a.bm + a.W;
    uint32_t* D2 = a.bm + 2 * size_t(a.W);
    uint32_t* D3 = a.bm + 3 * size_t(a.W);
    const uint32_t l0 = a.l[0], l1 = a.l[1], l2 = a.l[2], l3 = a.l[3];
    for (uint32_t r = wid; r < a.nroots; r += nw) {
        const uint32_t v = ldg(a.roots, r);
        if (T == 1) {  // edge: v = v0
            const Seg s1 = seg_of(a.rowptr, a.loff, a.L1, v, l1);
            if (s1.e > s1.b) {
                if (lane == 0) mark(D0, v);
                mark_seg(D1, a.lcol, s1, kNone, kNone);
            }
        } else if (T == 2) {  // wedge: v = v0 (center)
            const Seg s1 = seg_of(a.rowptr, a.loff, a.L1, v, l1), s2 = seg_of(a.rowptr, a.loff, a.L1, v, l2);
            if (s1.e > s1.b && s2.e - s2.b >= 1 + uint64_t(l1 == l2)) {
                if (lane == 0) mark(D0, v);
                mark_seg(D1, a.lcol, s1, kNone, kNone);
                mark_seg(D2, a.lcol, s2, kNone, kNone);
            }
        } else if (T == 3) {  // triangle: v = v0, lanes take v1
            const Seg A = seg_of(a.rowptr, a.loff, a.L1, v, l2), S1 = seg_of(a.rowptr, a.loff, a.L1, v, l1);
            bool any = false;
            if (A.e > A.b) {
                for (uint64_t k0 = S1.b; k0 < S1.e; k0 += 64) {
                    const uint64_t k = k0 + lane;
                    bool hit = false;
                    if (k < S1.e) {
                        const uint32_t v1 = ldg(a.lcol, k);
                        const Seg B = seg_of(a.rowptr, a.loff, a.L1, v1, l2);
                        // walk the shorter list, search the longer
                        const bool ab = A.e - A.b <= B.e - B.b;
                        const Seg sm = ab ? A : B, lg = ab ? B : A;
                        for (uint64_t q = sm.b; q < sm.e; ++q) {
                            const uint32_t w = ldg(a.lcol, q);
                            if (seg_has(a.lcol, lg, w)) {
                                mark(D2, w);
                                hit = true;
                            }
                        }
                        if (hit) mark(D1, v1);
                    }
                    if (__ballot(hit)) any = true;
                }
            }
            if (any && lane == 0) mark(D0, v);
        } else if (T == 4) {  // four-chain v2 - v0 - v1 - v3, from the v0 side: v = v0
            const Seg A2 = seg_of(a.rowptr, a.loff, a.L1, v, l2), S1 = seg_of(a.rowptr, a.loff, a.L1, v, l1);
            const uint64_t s2 = (A2.e - A2.b) - uint64_t(l1 == l2);
            bool any = false;
            uint32_t i1 = kNone, i2 = kNone;
            if (s2 >= 1) {
                for (uint64_t k0 = S1.b; k0 < S1.e; k0 += 64) {
                    const uint64_t k = k0 + lane;
                    bool ok = false;
                    uint32_t v1 = kNone, x2 = kNone;
                    if (k < S1.e) {
                        v1 = ldg(a.lcol, k);
                        const Seg B3 = seg_of(a.rowptr, a.loff, a.L1, v1, l3);
                        const uint64_t s3 = (B3.e - B3.b) - uint64_t(l0 == l3);
                        if (s3 >= 1) {
                            const uint32_t e3 = s3 == 1 ? only_other(a.lcol, B3, v) : kNone;
                            const uint32_t e2 = s2 == 1 ? only_other(a.lcol, A2, v1) : kNone;
                            ok = !(s2 == 1 && s3 == 1 && e2 == e3);
                            x2 = e3;
                        }
                        if (ok) mark(D1, v1);
                    }
                    fold_exclusions(ok, v1, x2, any, i1, i2);
                }
            }
            if (any) {
                if (lane == 0) mark(D0, v);
                mark_seg(D2, a.lcol, A2, i1, i2);
            }
        } else if (T == 5) {  // four-chain, from the v1 side: v = v1, lanes take v0
            const Seg B3 = seg_of(a.rowptr, a.loff, a.L1, v, l3), S0 = seg_of(a.rowptr, a.loff, a.L1, v, l0);
            const uint64_t s3 = (B3.e - B3.b) - uint64_t(l0 == l3);
            bool any = false;
            uint32_t j1 = kNone, j2 = kNone;
            if (s3 >= 1) {
                for (uint64_t k0 = S0.b; k0 < S0.e; k0 += 64) {
                    const uint64_t k = k0 + lane;
                    bool ok = false;
                    uint32_t v0 = kNone, y2 = kNone;
                    if (k < S0.e) {
                        v0 = ldg(a.lcol, k);
                        const Seg A2 = seg_of(a.rowptr, a.loff, a.L1, v0, l2);
                        const uint64_t s2 = (A2.e - A2.b) - uint64_t(l1 == l2);
                        if (s2 >= 1) {
                            const uint32_t e2 = s2 == 1 ? only_other(a.lcol, A2, v) : kNone;
                            const uint32_t e3 = s3 == 1 ? only_other(a.lcol, B3, v0) : kNone;
                            ok = !(s2 == 1 && s3 == 1 && e2 == e3);
                            y2 = e2;
                        }
                    }
                    fold_exclusions(ok, v0, y2, any, j1, j2);
                }
            }
            if (any) mark_seg(D3, a.lcol, B3, j1, j2);
        } else {  // three-star: v = v1 (center), leaves of labels l0, l2, l3
            const Seg S0 = seg_of(a.rowptr, a.loff, a.L1, v, l0), S2 = seg_of(a.rowptr, a.loff, a.L1, v, l2),
                      S3 = seg_of(a.rowptr, a.loff, a.L1, v, l3);
            const uint64_t r0 = 1 + uint64_t(l2 == l0) + uint64_t(l3 == l0);
            const uint64_t r2 = 1 + uint64_t(l0 == l2) + uint64_t(l3 == l2);
            const uint64_t r3 = 1 + uint64_t(l0 == l3) + uint64_t(l2 == l3);
            if (S0.e - S0.b >= r0 && S2.e - S2.b >= r2 && S3.e - S3.b >= r3) {
                if (lane == 0) mark(D1, v);
                mark_seg(D0, a.lcol, S0, kNone, kNone);
                mark_seg(D2, a.lcol, S2, kNone, kNone);
                mark_seg(D3, a.lcol, S3, kNone, kNone);
            }
        }
    }
}

// sizes[c * 4 + i] += popcount of bitmap i of candidate c (blockIdx.y = c * 4 + i).
__global__ __launch_bounds__(256) void fsm_popcount_kernel(const uint32_t* __restrict__ bm, uint32_t W,
                                                           uint32_t* __restrict__ sizes) {
    const uint32_t* b = bm + size_t(blockIdx.y) * W;
    uint32_t s = 0;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < W; i += gridDim.x * blockDim.x) s += __popc(b[i]);
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
    if ((threadIdx.x & 63) == 0 && s) atomicAdd(&sizes[blockIdx.y], s);
}

struct Pat {
    int size;
    uint32_t l[4];
    bool operator<(const Pat& o) const {
        if (size != o.size) return size < o.size;
        return std::lexicographical_compare(l, l + size, o.l, o.l + o.size);
    }
};
struct Cand {
    Pat canon;
    std::vector<Pat> autos;
};
Pat P2(uint32_t a, uint32_t b) { return Pat{2, {a, b, 0, 0}}; }
Pat P3(uint32_t a, uint32_t b, uint32_t c) { return Pat{3, {a, b, c, 0}}; }
Pat P4(uint32_t a, uint32_t b, uint32_t c, uint32_t d) { return Pat{4, {a, b, c, d}}; }

}  // namespace

// Uploads labels and builds the label-sorted lists and label vertex groups.
int fsm_set_labels(DeviceGraph& d, const uint32_t* h_lab) {
    if (int e = ensure_ready()) return e;
    if (d.oriented) KZ_FAIL(KZ_ERR_INVALID, "labels are for symmetric graphs (frequent subgraph mining)");
    auto F = std::make_shared<FsmState>();
    F->n = d.n;
    std::vector<uint32_t> cnt(KZ_MAX_LABELS + 1, 0);
    for (uint32_t v = 0; v < d.n; ++v) {
        if (h_lab[v] >= KZ_MAX_LABELS)
            KZ_FAIL(KZ_ERR_INVALID, "label %u of vertex %u is not below %d", h_lab[v], v, KZ_MAX_LABELS);
        F->L = std::max(F->L, h_lab[v] + 1);
        ++cnt[h_lab[v]];
    }
    F->voff.assign(F->L + 1, 0);
    for (uint32_t l = 0; l < F->L; ++l) F->voff[l + 1] = F->voff[l] + cnt[l];
    std::vector<uint32_t> vl(std::max<uint32_t>(1, d.n));
    {
        std::vector<uint32_t> at(F->voff.begin(), F->voff.end() - 1);
        for (uint32_t v = 0; v < d.n; ++v) vl[at[h_lab[v]]++] = v;
    }
    const size_t n1 = std::max<uint32_t>(1, d.n);
    const size_t m1 = std::max<uint64_t>(1, d.nnz);
    const size_t offs = size_t(F->L + 1) * n1;
    if (hipMalloc(&F->lab, 4 * n1) != hipSuccess || hipMalloc(&F->vlist, 4 * n1) != hipSuccess ||
        hipMalloc(&F->lcol, 4 * m1) != hipSuccess || hipMalloc(&F->loff, 4 * offs) != hipSuccess) {
        (void)hipGetLastError();
        KZ_FAIL(KZ_ERR_NOMEM, "labelled lists of %llu entries", (unsigned long long)d.nnz);
    }
    F->bytes = 8 * n1 + 4 * m1 + 4 * offs;
    hipStream_t s = rt().stream;
    if (d.n) {
        KZ_HIP(hipMemcpyAsync(F->lab, h_lab, 4 * size_t(d.n), hipMemcpyHostToDevice, s));
        KZ_HIP(hipMemcpyAsync(F->vlist, vl.data(), 4 * size_t(d.n), hipMemcpyHostToDevice, s));
        const uint32_t grid = std::min<uint32_t>((d.n + 3) / 4, uint32_t(rt().num_cus) * 16);
        hipLaunchKernelGGL(label_sort_kernel, dim3(grid), dim3(256), 0, s, d.rowptr, d.col, F->lab, d.n, F->L,
                           F->lcol, F->loff);
        KZ_HIP(hipGetLastError());
    }
    KZ_HIP(hipStreamSynchronize(s));
    d.fsm = F;
    return KZ_OK;
}

// The domain kernels of one candidate of a level (1 edge .. 5 three-star)
// with labels l[], into its 4 bitmaps bm[4 * W].
static void launch_domains(const DeviceGraph& d, const FsmState& F, int level, const uint32_t* l, uint32_t* bm,
                           uint32_t W, uint32_t grid, hipStream_t s) {
    FsmArgs a{d.rowptr, F.lcol, F.loff, F.L + 1, nullptr, 0, {0, 0, 0, 0}, bm, W};
    const int size = level == 1 ? 2 : (level <= 3 ? 3 : 4);
    for (int i = 0; i < size; ++i) a.l[i] = l[i];
    // roots: the vertices of the first label; the three-star and the
    // four-chain's second pass start from their center / v1
    auto roots = [&](uint32_t lab) {
        a.roots = F.vlist + F.voff[lab];
        a.nroots = F.voff[lab + 1] - F.voff[lab];
    };
    switch (level) {
        case 1:
            roots(l[0]);
            if (a.nroots) hipLaunchKernelGGL(fsm_domain_kernel<1>, dim3(grid), dim3(256), 0, s, a);
            break;
        case 2:
            roots(l[0]);
            if (a.nroots) hipLaunchKernelGGL(fsm_domain_kernel<2>, dim3(grid), dim3(256), 0, s, a);
            break;
        case 3:
            roots(l[0]);
            if (a.nroots) hipLaunchKernelGGL(fsm_domain_kernel<3>, dim3(grid), dim3(256), 0, s, a);
            break;
        case 4:
            roots(l[0]);
            if (a.nroots) hipLaunchKernelGGL(fsm_domain_kernel<4>, dim3(grid), dim3(256), 0, s, a);
            roots(l[1]);
            if (a.nroots) hipLaunchKernelGGL(fsm_domain_kernel<5>, dim3(grid), dim3(256), 0, s, a);
            break;
        default:
            roots(l[1]);
            if (a.nroots) hipLaunchKernelGGL(fsm_domain_kernel<6>, dim3(grid), dim3(256), 0, s, a);
            break;
    }
}

int fsm_domains(DeviceGraph& d, int level, const uint32_t* labels, uint32_t* sizes) {
    if (int e = ensure_ready()) return e;
    if (!d.fsm) KZ_FAIL(KZ_ERR_STATE, "the graph has no labels (kz_graph_set_labels)");
    if (level < KZ_FSM_EDGE || level > KZ_FSM_THREE_STAR) KZ_FAIL(KZ_ERR_INVALID, "level %d is not a pattern level", level);
    FsmState& F = *static_cast<FsmState*>(d.fsm.get());
    const int size = level == 1 ? 2 : (level <= 3 ? 3 : 4);
    for (int i = 0; i < size; ++i)
        if (labels[i] >= F.L) {
            // a label no vertex carries: every domain is empty
            for (int j = 0; j < 4; ++j) sizes[j] = 0;
            return KZ_OK;
        }
    hipStream_t s = rt().stream;
    const uint32_t W = std::max<uint32_t>(1, (F.n + 31) / 32);
    uint32_t *bm = nullptr, *sz = nullptr;
    if (hipMalloc(&bm, 16 * size_t(W)) != hipSuccess || hipMalloc(&sz, 16) != hipSuccess) {
        (void)hipGetLastError();
        hipFree(bm);
        KZ_FAIL(KZ_ERR_NOMEM, "domain bitmaps of %u vertices", F.n);
    }
    struct Free {
        uint32_t *a, *b;
        ~Free() {
            hipFree(a);
            hipFree(b);
        }
    } free_on_return{bm, sz};
    KZ_HIP(hipMemsetAsync(bm, 0, 16 * size_t(W), s));
    KZ_HIP(hipMemsetAsync(sz, 0, 16, s));
    launch_domains(d, F, level, labels, bm, W, uint32_t(rt().num_cus) * 8, s);
    const uint32_t pgrid = std::min<uint32_t>((W + 255) / 256, 64);
    hipLaunchKernelGGL(fsm_popcount_kernel, dim3(pgrid, 4), dim3(256), 0, s, bm, W, sz);
    KZ_HIP(hipGetLastError());
    KZ_HIP(hipMemcpyAsync(sizes, sz, 16, hipMemcpyDeviceToHost, s));
    KZ_HIP(hipStreamSynchronize(s));
    for (int j = size; j < 4; ++j) sizes[j] = 0;
    return KZ_OK;
}

int fsm_mine(DeviceGraph& d, uint32_t threshold, kz_fsm_stats* st, kz_fsm_candidate* out, uint32_t cap,
             uint32_t* ncand) {
    if (int e = ensure_ready()) return e;
    if (!d.fsm) KZ_FAIL(KZ_ERR_STATE, "the graph has no labels (kz_graph_set_labels)");
    FsmState& F = *static_cast<FsmState*>(d.fsm.get());
    const auto t0 = std::chrono::steady_clock::now();
    hipStream_t s = rt().stream;
    const uint32_t L = F.L, n = F.n;
    const uint32_t W = std::max<uint32_t>(1, (n + 31) / 32);
    std::memset(st, 0, sizeof(*st));
    st->num_labels = L;
    // frequent labels (setup, three_fsm.cc:255-275): every label's count
    // enters the level-0 support sum
    std::vector<char> fv(L);
    for (uint32_t l = 0; l < L; ++l) {
        const uint32_t c = F.voff[l + 1] - F.voff[l];
        fv[l] = c >= threshold;
        st->num_frequent[0] += fv[l];
        st->support_sum[0] += c;
    }
    std::set<Pat> fe, fw;  // frequent edges / wedges with their automorphisms
    uint32_t nc = 0;
    hipEvent_t e0, e1;
    KZ_HIP(hipEventCreate(&e0));
    KZ_HIP(hipEventCreate(&e1));
    float dev_ms = 0.f;
    // device buffers, grown per level
    Scratch bits, sizes;
    uint32_t* h_sizes = nullptr;
    KZ_HIP(hipHostMalloc(&h_sizes, 4 * 4 * 4096));
    auto release = [&]() {
        hipFree(bits.ptr);
        hipFree(sizes.ptr);
        hipHostFree(h_sizes);
        hipEventDestroy(e0);
        hipEventDestroy(e1);
    };
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) {
        (void)hipGetLastError();
        release();
        KZ_FAIL(KZ_ERR_DEVICE, "hipMemGetInfo failed");
    }
    // candidates per batch: 4 bitmaps each, at most a quarter of free memory
    const size_t per_cand = 16 * size_t(W);
    const uint32_t batch = uint32_t(std::max<size_t>(1, std::min<size_t>(4096, (free_b / 4) / per_cand)));
    const uint32_t grid = uint32_t(rt().num_cus) * 8;
    auto evaluate = [&](int level, const std::vector<Cand>& cs, std::set<Pat>* table) -> int {
        st->num_candidates[level] = uint32_t(cs.size());
        for (size_t c0 = 0; c0 < cs.size(); c0 += batch) {
            const uint32_t nb = uint32_t(std::min<size_t>(batch, cs.size() - c0));
            if (int e = scratch_reserve(bits, per_cand * nb)) return e;
            if (int e = scratch_reserve(sizes, 16 * size_t(nb))) return e;
            uint32_t* bm = static_cast<uint32_t*>(bits.ptr);
            uint32_t* sz = static_cast<uint32_t*>(sizes.ptr);
            KZ_HIP(hipEventRecord(e0, s));
            KZ_HIP(hipMemsetAsync(bm, 0, per_cand * nb, s));
            KZ_HIP(hipMemsetAsync(sz, 0, 16 * size_t(nb), s));
            for (uint32_t c = 0; c < nb; ++c) {
                const Pat& p = cs[c0 + c].canon;
                uint32_t l[4] = {0, 0, 0, 0};
                for (int i = 0; i < p.size; ++i) l[i] = p.l[i];
                launch_domains(d, F, level, l, bm + size_t(c) * 4 * W, W, grid, s);
            }
            const uint32_t pgrid = std::min<uint32_t>((W + 255) / 256, 64);
            hipLaunchKernelGGL(fsm_popcount_kernel, dim3(pgrid, 4 * nb), dim3(256), 0, s, bm, W, sz);
            KZ_HIP(hipGetLastError());
            KZ_HIP(hipEventRecord(e1, s));
            // sizes go through the pinned buffer in slices of 4096 candidates
            KZ_HIP(hipMemcpyAsync(h_sizes, sz, 16 * size_t(nb), hipMemcpyDeviceToHost, s));
            KZ_HIP(hipStreamSynchronize(s));
            float ms = 0.f;
            KZ_HIP(hipEventElapsedTime(&ms, e0, e1));
            dev_ms += ms;
            for (uint32_t c = 0; c < nb; ++c) {
                const Cand& cd = cs[c0 + c];
                uint32_t sup = n;  // get_support starts from the vertex count
                for (int i = 0; i < cd.canon.size; ++i) sup = std::min(sup, h_sizes[c * 4 + i]);
                const bool fr = sup >= threshold;
                if (out && nc < cap) {
                    kz_fsm_candidate& o = out[nc];
                    o.level = uint32_t(level);
                    for (int i = 0; i < 4; ++i) o.labels[i] = i < cd.canon.size ? cd.canon.l[i] : 0u;
                    o.support = sup;
                    o.frequent = fr;
                }
                ++nc;
                if (fr) {
                    if (table) {
                        table->insert(cd.canon);
                        for (const Pat& q : cd.autos) table->insert(q);
                    }
                    ++st->num_frequent[level];
                    st->support_sum[level] += sup;
                }
            }
        }
        return KZ_OK;
    };
    auto has = [](const std::set<Pat>& t, const Pat& p) { return t.count(p) != 0; };
    int err = KZ_OK;
    std::vector<Cand> cs;
    // edges (l0 <= l1), three_fsm.cc:282-304
    for (uint32_t a = 0; a < L; ++a)
        for (uint32_t b = a; b < L; ++b)
            if (fv[a] && fv[b]) cs.push_back({P2(a, b), {P2(b, a)}});
    if (!err) err = evaluate(1, cs, &fe);
    // wedges (center, leaf, leaf), :306-339
    cs.clear();
    for (uint32_t a = 0; a < L && !err; ++a) {
        if (!fv[a]) continue;
        for (uint32_t b = 0; b < L; ++b)
            if (has(fe, P2(a, b)))
                for (uint32_t c = b; c < L; ++c)
                    if (has(fe, P2(a, c))) cs.push_back({P3(a, b, c), {P3(a, c, b)}});
    }
    if (!err) err = evaluate(2, cs, &fw);
    // triangles (l0 <= l1 <= l2), :341-399; their table is never read again
    cs.clear();
    for (uint32_t a = 0; a < L && !err; ++a) {
        if (!fv[a]) continue;
        for (uint32_t b = a; b < L; ++b)
            if (has(fe, P2(a, b)))
                for (uint32_t c = b; c < L; ++c)
                    if (has(fw, P3(a, b, c)) && has(fw, P3(b, a, c)) && has(fw, P3(c, a, b)))
                        cs.push_back({P3(a, b, c), {}});
    }
    if (!err) err = evaluate(3, cs, nullptr);
    // four-chains v2 - v0 - v1 - v3 (l0 <= l1), :401-459
    cs.clear();
    for (uint32_t a = 0; a < L && !err; ++a) {
        if (!fv[a]) continue;
        for (uint32_t b = a; b < L; ++b) {
            if (!has(fe, P2(a, b))) continue;
            for (uint32_t c = 0; c < L; ++c) {
                if (!has(fw, P3(a, b, c)) || !has(fe, P2(a, c))) continue;
                for (uint32_t dd = 0; dd < L; ++dd)
                    if (has(fw, P3(b, a, dd)) && has(fe, P2(b, dd))) cs.push_back({P4(a, b, c, dd), {}});
            }
        }
    }
    if (!err) err = evaluate(4, cs, nullptr);
    // three-stars (leaf, center, leaf, leaf; leaves l1 <= l2 <= l3), :461-533
    cs.clear();
    for (uint32_t a = 0; a < L && !err; ++a) {
        if (!fv[a]) continue;
        for (uint32_t b = 0; b < L; ++b) {
            if (!has(fe, P2(a, b))) continue;
            for (uint32_t c = b; c < L; ++c) {
                if (!has(fw, P3(a, b, c))) continue;
                for (uint32_t dd = c; dd < L; ++dd)
                    if (has(fw, P3(a, c, dd)) && has(fw, P3(a, b, dd))) cs.push_back({P4(b, a, c, dd), {}});
            }
        }
    }
    if (!err) err = evaluate(5, cs, nullptr);
    release();
    if (err) return err;
    if (ncand) *ncand = nc;
    st->device_ms = dev_ms;
    st->total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return KZ_OK;
}

}  // namespace kz

using namespace kz;

extern "C" {

int kz_graph_set_labels(kz_graph g, const uint32_t* labels) {
    if (!g || (!labels && g->g.n)) KZ_FAIL(KZ_ERR_INVALID, "null argument");
    return fsm_set_labels(g->g, labels);
}

int kz_fsm_domains(kz_graph g, uint32_t level, const uint32_t* labels, uint32_t* sizes) {
    if (!g || !labels || !sizes) KZ_FAIL(KZ_ERR_INVALID, "null argument");
    return fsm_domains(g->g, int(level), labels, sizes);
}

int kz_fsm_mine(kz_graph g, uint32_t threshold, kz_fsm_stats* stats, kz_fsm_candidate* cands, uint32_t cap,
                uint32_t* ncand) {
    if (!g || !stats) KZ_FAIL(KZ_ERR_INVALID, "null argument");
    return fsm_mine(g->g, threshold, stats, cands, cap, ncand);
}

}  // extern "C"
