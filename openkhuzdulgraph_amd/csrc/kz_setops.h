// Wave-level set primitives shared by the plan executor and the batched
// set-operation entry point (src/vertex_set.cc:20-111 restated for wave64).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "kz_kernels.h"

namespace kz {

constexpr uint32_t kHashCap = 2048;  // per-wave LDS hash slots

struct SetView {
    const uint32_t* p;
    uint32_t n;
};

__device__ __forceinline__ uint64_t lanemask_lt() {
    const uint32_t lane = threadIdx.x & 63;
    return lane ? (~0ull >> (64 - lane)) : 0ull;
}

// Position class of v against sorted distinct matched values z[0..q).
__device__ __forceinline__ uint32_t pos_class(uint32_t v, const uint32_t* z, int q) {
    uint32_t lt = 0, eq = 0;
    for (int i = 0; i < q; ++i) {
        lt += z[i] < v;
        eq |= z[i] == v;
    }
    return 2 * lt + eq;
}

// What this wave's LDS hash currently holds (wave-uniform).  Only operands
// shared by sibling embeddings (an ancestor's N(v) or cached prefix, both
// immutable during a launch) are remembered, so consecutive siblings reuse
// the table instead of rebuilding it.
struct HashCache {
    const uint32_t* p;
    uint32_t n;
};

__device__ __forceinline__ void hash_geometry(uint32_t n, uint32_t* hmask, uint32_t* hshift) {
    uint32_t H = 64;
    while (H < 2 * n) H <<= 1;
    uint32_t bits = 0;
    while ((1u << bits) < H) ++bits;
    *hmask = H - 1;
    *hshift = 32 - bits;
}

// Wave-cooperative A op B (A & B or A \ B, src/vertex_set.cc:20-86).  The
// probed operand P is staged in this wave's LDS hash, the other operand S
// is streamed in 64-wide coalesced chunks; very skewed pairs stream the
// short list and binary-search the long one.  For every chunk `emit(x,
// keep, ballot, k)` receives each lane's element, whether it belongs to the
// result, the wave's ballot of `keep` and the number of results so far; the
// result order is S's order, i.e. ascending.  Returns the result length.
template <class Emit>
__device__ uint32_t wave_set_op(int op, SetView A, SetView B, bool a_shared, bool b_shared, uint32_t* hk,
                                HashCache& hc, Emit&& emit) {
    const uint32_t lane = threadIdx.x & 63;
    if (op == KZ_OP_INTERSECT && (A.n == 0 || B.n == 0)) return 0;
    if (op == KZ_OP_SUBTRACT && A.n == 0) return 0;
    constexpr uint32_t kFit = kHashCap / 2;
    SetView S, P;
    bool p_shared;
    bool search_mode;
    if (op == KZ_OP_SUBTRACT) {
        S = A;
        P = B;
        p_shared = b_shared;
        search_mode = P.n > kFit || (!p_shared && uint64_t(P.n) > 32ull * S.n);
    } else {
        // hash a sibling-shared operand when it fits, else the shorter one
        if (a_shared && A.n <= kFit && !(b_shared && B.n < A.n)) {
            P = A;
            S = B;
            p_shared = true;
        } else if (b_shared && B.n <= kFit) {
            P = B;
            S = A;
            p_shared = true;
        } else if (A.n <= B.n) {
            P = A;
            S = B;
            p_shared = a_shared;
        } else {
            P = B;
            S = A;
            p_shared = b_shared;
        }
        const bool cached = p_shared && hc.p == P.p && hc.n == P.n;
        if (!cached && (P.n > kFit || uint64_t(S.n) > 32ull * P.n)) {
            // stream the shorter list, binary-search the longer one
            if (P.n <= S.n) {
                SetView t = S;
                S = P;
                P = t;
            }
            search_mode = true;
        } else {
            search_mode = false;
        }
    }
    uint32_t hmask = 0, hshift = 0;
    if (!search_mode && P.n) {
        hash_geometry(P.n, &hmask, &hshift);
        if (!(p_shared && hc.p == P.p && hc.n == P.n)) {
            for (uint32_t i = lane; i <= hmask; i += 64) hk[i] = kEmpty;
            wave_lds_sync();
            for (uint32_t i = lane; i < P.n; i += 64) {
                const uint32_t x = P.p[i];
                uint32_t h = (x * 2654435761u) >> hshift;
                while (atomicCAS(&hk[h], kEmpty, x) != kEmpty) h = (h + 1) & hmask;
            }
            wave_lds_sync();
            hc.p = p_shared ? P.p : nullptr;
            hc.n = p_shared ? P.n : 0;
        }
    }
    uint32_t k = 0;
    for (uint32_t base = 0; base < S.n; base += 64) {
        const uint32_t i = base + lane;
        const bool valid = i < S.n;
        const uint32_t x = valid ? S.p[i] : 0;
        bool found = false;
        if (valid) {
            if (search_mode) {
                found = bsearch_global(P.p, P.n, x);
            } else if (P.n) {
                uint32_t h = (x * 2654435761u) >> hshift;
                while (true) {
                    const uint32_t key = hk[h];
                    if (key == x) {
                        found = true;
                        break;
                    }
                    if (key == kEmpty) break;
                    h = (h + 1) & hmask;
                }
            }
        }
        const bool keep = valid && (op == KZ_OP_INTERSECT ? found : !found);
        const uint64_t m = __ballot(keep);
        emit(x, keep, m, k);
        k += __popcll(m);
    }
    return k;
}

}  // namespace kz
