// Batched set primitives: the VertexSet operations of src/vertex_set.cc and
// the search helper of include/utilities.h:76-86 applied to many list pairs
// in one launch (one wave64 per pair, persistent waves).
//   intersect_with        (vertex_set.cc:20-49)  -> sorted A & B
//   subtracted_by         (vertex_set.cc:51-86)  -> sorted A \ B
//   intersect_and_count   (vertex_set.cc:88-111) -> |A & B|
//   trim_v_list_lower_bound (utilities.h:76-86)  -> index of first id > bound

#include <algorithm>
#include <cstring>

#include "kz_device.h"
#include "kz_kernels.h"
#include "kz_setops.h"

namespace kz {

constexpr int kSetWaves = 4;

__global__ __launch_bounds__(256) void set_op_batch_kernel(int op, const uint32_t* __restrict__ values,
                                                           const uint64_t* __restrict__ list_off,
                                                           const uint32_t* __restrict__ a_idx,
                                                           const uint32_t* __restrict__ b_idx,
                                                           const uint32_t* __restrict__ bounds, uint32_t num_pairs,
                                                           uint32_t* __restrict__ out_values,
                                                           const uint64_t* __restrict__ out_off,
                                                           uint32_t* __restrict__ out_len) {
    __shared__ uint32_t sHash[kSetWaves][kHashCap];
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t w = threadIdx.x >> 6;
    uint32_t* hk = sHash[w];
    HashCache hc{nullptr, 0};
    const uint32_t nw = gridDim.x * kSetWaves;
    for (uint32_t i = blockIdx.x * kSetWaves + w; i < num_pairs; i += nw) {
        const uint32_t ia = a_idx[i];
        SetView A{values + list_off[ia], uint32_t(list_off[ia + 1] - list_off[ia])};
        if (op == KZ_SETOP_TRIM_LOWER_BOUND) {
            if (lane == 0) out_len[i] = upper_bound_global(A.p, A.n, bounds[i]);
            continue;
        }
        const uint32_t ib = b_idx[i];
        SetView B{values + list_off[ib], uint32_t(list_off[ib + 1] - list_off[ib])};
        uint32_t n;
        if (op == KZ_SETOP_INTERSECT_COUNT) {
            n = wave_set_op(KZ_OP_INTERSECT, A, B, false, false, hk, hc, [&](uint32_t, bool, uint64_t, uint32_t) {});
        } else {
            uint32_t* outp = out_values + out_off[i];
            n = wave_set_op(op == KZ_SETOP_INTERSECT ? KZ_OP_INTERSECT : KZ_OP_SUBTRACT, A, B, false, false, hk, hc,
                            [&](uint32_t x, bool keep, uint64_t msk, uint32_t k) {
                                if (keep) outp[k + __popcll(msk & lanemask_lt())] = x;
                            });
        }
        if (lane == 0) out_len[i] = n;
    }
}

}  // namespace kz

using namespace kz;

extern "C" int kz_set_op_batch(int op, const uint32_t* values, const uint64_t* list_off, uint32_t num_lists,
                               const uint32_t* a_idx, const uint32_t* b_idx, const uint32_t* bounds,
                               uint32_t num_pairs, uint32_t* out_values, const uint64_t* out_off, uint32_t* out_len) {
    if (op < KZ_SETOP_INTERSECT || op > KZ_SETOP_TRIM_LOWER_BOUND) KZ_FAIL(KZ_ERR_INVALID, "unknown set op %d", op);
    if (!list_off || !a_idx || !out_len || (num_lists && !values)) KZ_FAIL(KZ_ERR_INVALID, "null argument");
    const bool writes = op == KZ_SETOP_INTERSECT || op == KZ_SETOP_SUBTRACT;
    if (op != KZ_SETOP_TRIM_LOWER_BOUND && !b_idx) KZ_FAIL(KZ_ERR_INVALID, "b_idx required");
    if (op == KZ_SETOP_TRIM_LOWER_BOUND && !bounds) KZ_FAIL(KZ_ERR_INVALID, "bounds required");
    if (writes && (!out_off || !out_values)) KZ_FAIL(KZ_ERR_INVALID, "out_off and out_values required");
    if (int e = ensure_ready()) return e;
    if (num_pairs == 0) return KZ_OK;
    const uint64_t nvals = list_off[num_lists];
    for (uint32_t i = 0; i < num_pairs; ++i) {
        if (a_idx[i] >= num_lists || (op != KZ_SETOP_TRIM_LOWER_BOUND && b_idx[i] >= num_lists))
            KZ_FAIL(KZ_ERR_INVALID, "pair %u names a list out of range", i);
        if (writes) {
            const uint64_t na = list_off[a_idx[i] + 1] - list_off[a_idx[i]];
            const uint64_t nb = list_off[b_idx[i] + 1] - list_off[b_idx[i]];
            const uint64_t need = op == KZ_SETOP_INTERSECT ? std::min(na, nb) : na;
            if (out_off[i + 1] - out_off[i] < need) KZ_FAIL(KZ_ERR_INVALID, "output slot %u too small", i);
        }
    }
    hipStream_t s = rt().stream;
    const uint64_t nout = writes ? out_off[num_pairs] : 0;
    uint32_t *d_vals = nullptr, *d_a = nullptr, *d_b = nullptr, *d_bd = nullptr, *d_out = nullptr, *d_len = nullptr;
    uint64_t *d_loff = nullptr, *d_ooff = nullptr;
    auto cleanup = [&]() {
        hipFree(d_vals);
        hipFree(d_a);
        hipFree(d_b);
        hipFree(d_bd);
        hipFree(d_out);
        hipFree(d_len);
        hipFree(d_loff);
        hipFree(d_ooff);
    };
    hipError_t err = hipSuccess;
    auto H = [&](hipError_t e) {
        if (err == hipSuccess) err = e;
    };
    H(hipMalloc(&d_vals, 4 * std::max<uint64_t>(1, nvals)));
    H(hipMalloc(&d_loff, 8 * (uint64_t(num_lists) + 1)));
    H(hipMalloc(&d_a, 4 * uint64_t(num_pairs)));
    H(hipMalloc(&d_b, 4 * uint64_t(num_pairs)));
    H(hipMalloc(&d_bd, 4 * uint64_t(num_pairs)));
    H(hipMalloc(&d_len, 4 * uint64_t(num_pairs)));
    H(hipMalloc(&d_ooff, 8 * (uint64_t(num_pairs) + 1)));
    H(hipMalloc(&d_out, 4 * std::max<uint64_t>(1, nout)));
    if (err != hipSuccess) {
        cleanup();
        (void)hipGetLastError();
        KZ_FAIL(KZ_ERR_NOMEM, "set op buffers");
    }
    if (nvals) H(hipMemcpyAsync(d_vals, values, 4 * nvals, hipMemcpyHostToDevice, s));
    H(hipMemcpyAsync(d_loff, list_off, 8 * (uint64_t(num_lists) + 1), hipMemcpyHostToDevice, s));
    H(hipMemcpyAsync(d_a, a_idx, 4 * uint64_t(num_pairs), hipMemcpyHostToDevice, s));
    if (b_idx) H(hipMemcpyAsync(d_b, b_idx, 4 * uint64_t(num_pairs), hipMemcpyHostToDevice, s));
    if (bounds) H(hipMemcpyAsync(d_bd, bounds, 4 * uint64_t(num_pairs), hipMemcpyHostToDevice, s));
    if (writes) H(hipMemcpyAsync(d_ooff, out_off, 8 * (uint64_t(num_pairs) + 1), hipMemcpyHostToDevice, s));
    const uint32_t grid = std::min<uint32_t>((num_pairs + kSetWaves - 1) / kSetWaves, uint32_t(rt().num_cus) * 8);
    hipLaunchKernelGGL(set_op_batch_kernel, dim3(grid), dim3(64 * kSetWaves), 0, s, op, d_vals, d_loff, d_a, d_b, d_bd,
                       num_pairs, d_out, d_ooff, d_len);
    H(hipGetLastError());
    H(hipMemcpyAsync(out_len, d_len, 4 * uint64_t(num_pairs), hipMemcpyDeviceToHost, s));
    if (writes && nout) H(hipMemcpyAsync(out_values, d_out, 4 * nout, hipMemcpyDeviceToHost, s));
    H(hipStreamSynchronize(s));
    cleanup();
    if (err != hipSuccess) KZ_FAIL(KZ_ERR_DEVICE, "set op batch: %s", hipGetErrorString(err));
    return KZ_OK;
}
