// Exact merge-model byte terms of the clique schedules (measurement only;
// the reference reports no bandwidth for this path).

#include <algorithm>

#include "kz_device.h"
#include "kz_kernels.h"

namespace kz {

// ------------------------------------------------------- workload terms --
// Exact terms of the merge-model byte count of the clique schedules (the
// kautomine four_clique.cc schedule reads N+(u), N+(v) per edge and the
// cached prefix c(u,v) plus N+(w) per triangle).  One workgroup per root u:
// N+(u) in an LDS hash, one wave per v in N+(u) streams N+(v).
//   acc[0] = sum_{(u,v)} c(u,v)            (triangles T)
//   acc[1] = sum_{(u,v)} c(u,v)^2
//   acc[2] = sum_{(u,v,w) in Tri} d+(w)
//   acc[3] = sum_{(u,v)} d+(v)             (= sum_v d-(v) d+(v))
__global__ __launch_bounds__(256) void workload_kernel(const uint64_t* __restrict__ rowptr,
                                                       const uint32_t* __restrict__ col, uint32_t n,
                                                       uint32_t* __restrict__ cursor, uint32_t H, uint32_t hshift,
                                                       unsigned long long* __restrict__ out) {
    extern __shared__ __align__(16) unsigned char smem[];
    uint32_t* hk = reinterpret_cast<uint32_t*>(smem);
    uint16_t* hv = reinterpret_cast<uint16_t*>(smem + size_t(H) * 4);
    __shared__ uint32_t sTask;
    __shared__ unsigned long long sRed[4][4];
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    unsigned long long a0 = 0, a1 = 0, a2 = 0, a3 = 0;
    TaskChunk chunk(n, gridDim.x, 8);
    while (true) {
        if (threadIdx.x == 0) sTask = thread_next_task(chunk, cursor);
        __syncthreads();
        const uint32_t u = sTask;
        if (u >= n) break;
        const uint64_t ub = rowptr[u];
        const uint32_t d = uint32_t(rowptr[u + 1] - ub);
        if (d > 0) {
            for (uint32_t i = threadIdx.x; i < H; i += blockDim.x) hk[i] = kEmpty;
            __syncthreads();
            for (uint32_t i = threadIdx.x; i < d; i += blockDim.x)
                hash_insert(hk, hv, H - 1, hshift, col[ub + i], uint16_t(i & 0xFFFF));
            __syncthreads();
            for (uint32_t i = wave; i < d; i += blockDim.x / 64) {
                const uint32_t v = col[ub + i];
                const uint64_t vb = rowptr[v];
                const uint32_t dv = uint32_t(rowptr[v + 1] - vb);
                uint32_t c = 0;
                unsigned long long sw = 0;
                for (uint32_t base = 0; base < dv; base += 64) {
                    const uint32_t q = base + lane;
                    if (q < dv) {
                        const uint32_t w = col[vb + q];
                        if (hash_find(hk, hv, H - 1, hshift, w) != kNotFound) {
                            c += 1;
                            sw += uint32_t(rowptr[w + 1] - rowptr[w]);
                        }
                    }
                }
                unsigned long long cc = wave_sum_u64(c);
                sw = wave_sum_u64(sw);
                if (lane == 0) {
                    a0 += cc;
                    a1 += cc * cc;
                    a2 += sw;
                    a3 += dv;
                }
            }
        }
        __syncthreads();
    }
    if (lane == 0) {
        sRed[wave][0] = a0;
        sRed[wave][1] = a1;
        sRed[wave][2] = a2;
        sRed[wave][3] = a3;
    }
    __syncthreads();
    if (threadIdx.x < 4) {
        unsigned long long s = 0;
        for (int w = 0; w < 4; ++w) s += sRed[w][threadIdx.x];
        atomicAdd(&out[threadIdx.x], s);
    }
}

int clique_workload(const CliqueView& gv, uint64_t* out6) {
    Runtime& R = rt();
    const uint32_t n = uint32_t(gv.h_fdeg.size());
    uint32_t maxd = 1;
    uint64_t m = 0, sq = 0;
    for (uint32_t v = 0; v < n; ++v) {
        maxd = std::max(maxd, gv.h_fdeg[v]);
        m += gv.h_fdeg[v];
        sq += uint64_t(gv.h_fdeg[v]) * gv.h_fdeg[v];
    }
    if (maxd > kMaxBitmapDegree)
        KZ_FAIL(KZ_ERR_UNSUPPORTED, "workload pass needs forward out-degree <= %u", kMaxBitmapDegree);
    uint32_t H = 64;
    while (H < 2 * maxd) H <<= 1;
    uint32_t bits = 0;
    while ((1u << bits) < H) ++bits;
    const size_t lds = size_t(H) * 6;
    uint32_t* cur = nullptr;
    unsigned long long* acc = nullptr;
    KZ_HIP(hipMalloc(&cur, 4));
    KZ_HIP(hipMalloc(&acc, 32));
    KZ_HIP(hipMemsetAsync(cur, 0, 4, R.stream));
    KZ_HIP(hipMemsetAsync(acc, 0, 32, R.stream));
    if (hipFuncSetAttribute(reinterpret_cast<const void*>(&workload_kernel), hipFuncAttributeMaxDynamicSharedMemorySize,
                            int(R.lds_per_block) - kStaticLdsReserve) != hipSuccess)
        (void)hipGetLastError();
    const uint32_t grid = uint32_t(R.num_cus) * std::max<uint32_t>(1, std::min<uint32_t>(8, uint32_t(R.lds_per_block / (lds + kStaticLdsReserve))));
    hipLaunchKernelGGL(workload_kernel, dim3(grid), dim3(256), lds, R.stream, gv.rowptr, gv.col, n, cur, H, 32 - bits,
                       acc);
    KZ_HIP(hipGetLastError());
    unsigned long long h[4];
    KZ_HIP(hipMemcpyAsync(h, acc, 32, hipMemcpyDeviceToHost, R.stream));
    KZ_HIP(hipStreamSynchronize(R.stream));
    hipFree(cur);
    hipFree(acc);
    out6[0] = m;
    out6[1] = h[0];
    out6[2] = h[1];
    out6[3] = h[2];
    out6[4] = sq;
    out6[5] = h[3];
    return KZ_OK;
}

}  // namespace kz
