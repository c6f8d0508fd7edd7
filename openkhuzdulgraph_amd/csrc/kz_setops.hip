// Batched set primitives: the VertexSet operations of src/vertex_set.cc and
// the search helper of include/utilities.h:76-86 applied to many list pairs
// in one launch (one wave64 per pair, persistent waves).
//   intersect_with        (vertex_set.cc:20-49)  -> sorted A & B
//   subtracted_by         (vertex_set.cc:51-86)  -> sorted A \ B
//   intersect_and_count   (vertex_set.cc:88-111) -> |A & B|
//   trim_v_list_lower_bound (utilities.h:76-86)  -> index of first id > bound

#include <algorithm>
#include <cstring>
#include <mutex>

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

// Device staging of one batch: inputs copied in, outputs copied back.
struct SetOpStage {
    uint32_t *vals, *a, *b, *bd, *out, *len;
    uint64_t *loff, *ooff;
};

// Byte offsets of the staging regions in one buffer, each 256-byte aligned.
struct SetOpLayout {
    size_t vals, loff, a, b, bd, len, ooff, out, total;
};

static size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

static SetOpLayout setop_layout(uint64_t nvals, uint32_t num_lists, uint32_t num_pairs, uint64_t nout) {
    SetOpLayout L;
    size_t at = 0;
    auto take = [&](size_t bytes) {
        const size_t o = at;
        at = align256(at + bytes);
        return o;
    };
    L.vals = take(4 * std::max<uint64_t>(1, nvals));
    L.loff = take(8 * (uint64_t(num_lists) + 1));
    L.a = take(4 * uint64_t(num_pairs));
    L.b = take(4 * uint64_t(num_pairs));
    L.bd = take(4 * uint64_t(num_pairs));
    L.len = take(4 * uint64_t(num_pairs));
    L.ooff = take(8 * (uint64_t(num_pairs) + 1));
    L.out = take(4 * std::max<uint64_t>(1, nout));
    L.total = at;
    return L;
}

static SetOpStage setop_stage(void* base, const SetOpLayout& L) {
    char* b = static_cast<char*>(base);
    return SetOpStage{reinterpret_cast<uint32_t*>(b + L.vals), reinterpret_cast<uint32_t*>(b + L.a),
                      reinterpret_cast<uint32_t*>(b + L.b),    reinterpret_cast<uint32_t*>(b + L.bd),
                      reinterpret_cast<uint32_t*>(b + L.out),  reinterpret_cast<uint32_t*>(b + L.len),
                      reinterpret_cast<uint64_t*>(b + L.loff), reinterpret_cast<uint64_t*>(b + L.ooff)};
}

// Argument checks shared by both entry points; sets the staged sizes.
static int setop_check(int op, const uint32_t* values, const uint64_t* list_off, uint32_t num_lists,
                       const uint32_t* a_idx, const uint32_t* b_idx, const uint32_t* bounds, uint32_t num_pairs,
                       uint32_t* out_values, const uint64_t* out_off, uint32_t* out_len, uint64_t* nvals,
                       uint64_t* nout) {
    if (op < KZ_SETOP_INTERSECT || op > KZ_SETOP_TRIM_LOWER_BOUND) KZ_FAIL(KZ_ERR_INVALID, "unknown set op %d", op);
    if (!list_off || !a_idx || !out_len || (num_lists && !values)) KZ_FAIL(KZ_ERR_INVALID, "null argument");
    const bool writes = op == KZ_SETOP_INTERSECT || op == KZ_SETOP_SUBTRACT;
    if (op != KZ_SETOP_TRIM_LOWER_BOUND && !b_idx) KZ_FAIL(KZ_ERR_INVALID, "b_idx required");
    if (op == KZ_SETOP_TRIM_LOWER_BOUND && !bounds) KZ_FAIL(KZ_ERR_INVALID, "bounds required");
    if (writes && (!out_off || !out_values)) KZ_FAIL(KZ_ERR_INVALID, "out_off and out_values required");
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
    *nvals = list_off[num_lists];
    *nout = writes && num_pairs ? out_off[num_pairs] : 0;
    return KZ_OK;
}

// Copies the batch into the staging regions, runs the kernel and copies the
// results back, on the library stream; returns after the stream is idle.
static int setop_run(const SetOpStage& d, int op, const uint32_t* values, const uint64_t* list_off,
                     uint32_t num_lists, const uint32_t* a_idx, const uint32_t* b_idx, const uint32_t* bounds,
                     uint32_t num_pairs, uint32_t* out_values, const uint64_t* out_off, uint32_t* out_len,
                     uint64_t nvals, uint64_t nout) {
    hipStream_t s = rt().stream;
    const bool writes = op == KZ_SETOP_INTERSECT || op == KZ_SETOP_SUBTRACT;
    hipError_t err = hipSuccess;
    auto H = [&](hipError_t e) {
        if (err == hipSuccess) err = e;
    };
    if (nvals) H(hipMemcpyAsync(d.vals, values, 4 * nvals, hipMemcpyHostToDevice, s));
    H(hipMemcpyAsync(d.loff, list_off, 8 * (uint64_t(num_lists) + 1), hipMemcpyHostToDevice, s));
    H(hipMemcpyAsync(d.a, a_idx, 4 * uint64_t(num_pairs), hipMemcpyHostToDevice, s));
    if (b_idx) H(hipMemcpyAsync(d.b, b_idx, 4 * uint64_t(num_pairs), hipMemcpyHostToDevice, s));
    if (bounds) H(hipMemcpyAsync(d.bd, bounds, 4 * uint64_t(num_pairs), hipMemcpyHostToDevice, s));
    if (writes) H(hipMemcpyAsync(d.ooff, out_off, 8 * (uint64_t(num_pairs) + 1), hipMemcpyHostToDevice, s));
    const uint32_t grid = std::min<uint32_t>((num_pairs + kSetWaves - 1) / kSetWaves, uint32_t(rt().num_cus) * 8);
    hipLaunchKernelGGL(set_op_batch_kernel, dim3(grid), dim3(64 * kSetWaves), 0, s, op, d.vals, d.loff, d.a, d.b,
                       d.bd, num_pairs, d.out, d.ooff, d.len);
    H(hipGetLastError());
    H(hipMemcpyAsync(out_len, d.len, 4 * uint64_t(num_pairs), hipMemcpyDeviceToHost, s));
    if (writes && nout) H(hipMemcpyAsync(out_values, d.out, 4 * nout, hipMemcpyDeviceToHost, s));
    H(hipStreamSynchronize(s));
    if (err != hipSuccess) {
        (void)hipGetLastError();
        KZ_FAIL(KZ_ERR_DEVICE, "set op batch: %s", hipGetErrorString(err));
    }
    return KZ_OK;
}

// Staging of kz_set_op_batch_noalloc: a file-scope device array (each device
// the library runs on holds its own copy until the process ends).  Host
// code uses only the address hipGetSymbolAddress returns for the device,
// never the array's name.
__device__ __align__(256) unsigned char scratch[KZ_SETOP_NOALLOC_MAX_BYTES];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;  // serializes the calls that share a device's array
void* g_scratch_ptr[kMaxDevices] = {};

// Address of `scratch` on the library's device, looked up once per device
// with that device current.
static int scratch_address(void** out) {
    const int dev = rt().device;
    if (dev < 0 || dev >= kMaxDevices) KZ_FAIL(KZ_ERR_UNSUPPORTED, "device %d beyond the scratch table", dev);
    if (!g_scratch_ptr[dev]) {
        int cur = -1;
        KZ_HIP(hipGetDevice(&cur));
        if (cur != dev) KZ_HIP(hipSetDevice(dev));
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (cur != dev) (void)hipSetDevice(cur);
        if (e != hipSuccess || !p) {
            (void)hipGetLastError();
            KZ_FAIL(KZ_ERR_DEVICE, "scratch address: %s", hipGetErrorString(e));
        }
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return KZ_OK;
}

}  // namespace kz

using namespace kz;

extern "C" int kz_set_op_batch(int op, const uint32_t* values, const uint64_t* list_off, uint32_t num_lists,
                               const uint32_t* a_idx, const uint32_t* b_idx, const uint32_t* bounds,
                               uint32_t num_pairs, uint32_t* out_values, const uint64_t* out_off, uint32_t* out_len) {
    uint64_t nvals = 0, nout = 0;
    if (int e = setop_check(op, values, list_off, num_lists, a_idx, b_idx, bounds, num_pairs, out_values, out_off,
                            out_len, &nvals, &nout))
        return e;
    if (int e = ensure_ready()) return e;
    if (num_pairs == 0) return KZ_OK;
    const SetOpLayout L = setop_layout(nvals, num_lists, num_pairs, nout);
    void* buf = nullptr;
    if (hipMalloc(&buf, L.total) != hipSuccess) {
        (void)hipGetLastError();
        KZ_FAIL(KZ_ERR_NOMEM, "set op buffers (%zu bytes)", L.total);
    }
    const int e = setop_run(setop_stage(buf, L), op, values, list_off, num_lists, a_idx, b_idx, bounds, num_pairs,
                            out_values, out_off, out_len, nvals, nout);
    (void)hipFree(buf);
    return e;
}

extern "C" int kz_set_op_batch_noalloc(int op, const uint32_t* values, const uint64_t* list_off, uint32_t num_lists,
                                       const uint32_t* a_idx, const uint32_t* b_idx, const uint32_t* bounds,
                                       uint32_t num_pairs, uint32_t* out_values, const uint64_t* out_off,
                                       uint32_t* out_len) {
    uint64_t nvals = 0, nout = 0;
    if (int e = setop_check(op, values, list_off, num_lists, a_idx, b_idx, bounds, num_pairs, out_values, out_off,
                            out_len, &nvals, &nout))
        return e;
    if (int e = ensure_ready()) return e;
    if (num_pairs == 0) return KZ_OK;
    const SetOpLayout L = setop_layout(nvals, num_lists, num_pairs, nout);
    if (L.total > KZ_SETOP_NOALLOC_MAX_BYTES)
        KZ_FAIL(KZ_ERR_INVALID, "batch needs %zu bytes of staging, more than the %llu of kz_set_op_batch_noalloc",
                L.total, (unsigned long long)KZ_SETOP_NOALLOC_MAX_BYTES);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    void* base = nullptr;
    if (int e = scratch_address(&base)) return e;
    return setop_run(setop_stage(base, L), op, values, list_off, num_lists, a_idx, b_idx, bounds, num_pairs,
                     out_values, out_off, out_len, nvals, nout);
}
