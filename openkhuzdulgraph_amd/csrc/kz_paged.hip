// Graphs beyond HBM (kz_paged_*): a degree-oriented CSR whose neighbour
// lists stay in pinned host memory, counted in batches of consecutive roots
// whose lists are fetched into HBM on demand.
//
// Reference: the remote graph-data path of the distributed engine,
// src/engine.cc:94-510 (GraphDataRequestSender: the neighbour lists an
// embedding batch needs are requested once per vertex, deduplicated, and
// served from a cache of high-degree vertices' lists) and its handler
// (:513-647, the owner sends the lists); the cache is sized by
// cache_degree_threshold and relative_cache_size
// (include/distributed_application.h:38-46).  On one GPU the "remote"
// store is host memory over PCIe:
//   * the row offsets and degrees stay in HBM (about 36 bytes per vertex
//     with the working arrays), the lists in hipHostMalloc'd memory that
//     kernels read directly;
//   * the lists of the highest-degree vertices (total degree >= the
//     threshold, largest first, as many as cache_bytes holds) are kept in
//     an HBM cache;
//   * roots are cut into batches of consecutive ids whose undeduplicated
//     list entries, d+(u) + sum over w in N+(u) of d+(w), fit the working
//     buffer (list_bytes);
//   * per batch: a request kernel stamps every vertex the batch touches
//     (the roots and their forward neighbours) with the batch number -- the
//     deduplication --, the stamped vertices are compacted into a request
//     list and given rows of a working CSR (a scan of the stamped degrees),
//     a gather kernel copies each requested list from the cache or from
//     host memory into it, and the persistent clique kernels count the
//     batch's roots on the working CSR.

#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <numeric>
#include <omp.h>
#include <vector>

#include "kz_device.h"
#include "kz_kernels.h"

namespace kz {
namespace {

constexpr uint64_t kNotCached = ~uint64_t(0);

struct StampedDeg {  // degree of a vertex stamped by this batch, else 0
    const uint32_t* stamp;
    const uint32_t* deg;
    uint32_t epoch;
    uint32_t n;
    __host__ __device__ uint64_t operator()(uint32_t v) const {
        return v < n && stamp[v] == epoch ? uint64_t(deg[v]) : 0ull;
    }
};

struct IsStamped {
    const uint32_t* stamp;
    uint32_t epoch;
    __host__ __device__ bool operator()(uint32_t v) const { return stamp[v] == epoch; }
};

// List of v: in the cache when coff[v] != kNotCached, else in host memory.
__device__ __forceinline__ const uint32_t* list_of(const uint64_t* __restrict__ rowptr,
                                                   const uint32_t* __restrict__ hcol,
                                                   const uint64_t* __restrict__ coff,
                                                   const uint32_t* __restrict__ ccol, uint32_t v) {
    const uint64_t c = coff[v];
    return c != kNotCached ? ccol + c : hcol + rowptr[v];
}

// Wave per root u in [r0, r1): stamps u and its forward neighbours.
__global__ __launch_bounds__(256) void request_kernel(uint32_t r0, uint32_t r1, const uint64_t* __restrict__ rowptr,
                                                      const uint32_t* __restrict__ deg,
                                                      const uint32_t* __restrict__ hcol,
                                                      const uint64_t* __restrict__ coff,
                                                      const uint32_t* __restrict__ ccol, uint32_t* __restrict__ stamp,
                                                      uint32_t epoch) {
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t nw = (gridDim.x * blockDim.x) >> 6;
    for (uint32_t u = r0 + ((blockIdx.x * blockDim.x + threadIdx.x) >> 6); u < r1; u += nw) {
        if (lane == 0) stamp[u] = epoch;
        const uint32_t d = deg[u];
        const uint32_t* L = list_of(rowptr, hcol, coff, ccol, u);
        for (uint32_t i = lane; i < d; i += 64) stamp[L[i]] = epoch;
    }
}

// Wave per requested vertex: copies its list into the working CSR and
// counts the bytes read from the cache and from host memory.
__global__ __launch_bounds__(256) void gather_kernel(const uint32_t* __restrict__ req, const uint32_t* __restrict__ nreq,
                                                     const uint64_t* __restrict__ rowptr,
                                                     const uint32_t* __restrict__ deg,
                                                     const uint32_t* __restrict__ hcol,
                                                     const uint64_t* __restrict__ coff,
                                                     const uint32_t* __restrict__ ccol,
                                                     const uint64_t* __restrict__ wrow, uint32_t* __restrict__ wcol,
                                                     unsigned long long* __restrict__ bytes) {
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t nw = (gridDim.x * blockDim.x) >> 6;
    const uint32_t total = *nreq;
    unsigned long long from_host = 0, from_cache = 0;
    for (uint32_t i = (blockIdx.x * blockDim.x + threadIdx.x) >> 6; i < total; i += nw) {
        const uint32_t v = req[i];
        const uint32_t d = deg[v];
        const uint64_t c = coff[v];
        const uint32_t* src = c != kNotCached ? ccol + c : hcol + rowptr[v];
        uint32_t* dst = wcol + wrow[v];
        for (uint32_t k = lane; k < d; k += 64) dst[k] = src[k];
        if (c != kNotCached) from_cache += 4ull * d;
        else from_host += 4ull * d;
    }
    if (lane == 0) {
        if (from_host) atomicAdd(&bytes[0], from_host);
        if (from_cache) atomicAdd(&bytes[1], from_cache);
    }
}

__global__ void fill_u32(uint32_t* __restrict__ p, uint64_t n, uint32_t v) {
    for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += uint64_t(gridDim.x) * blockDim.x)
        p[i] = v;
}

}  // namespace
}  // namespace kz

struct kz_paged_s {
    uint32_t n = 0;
    uint64_t m = 0;
    std::vector<uint32_t> h_deg;       // forward out-degrees
    std::vector<uint32_t> batch_end;   // root batches [end[i-1], end[i])
    std::vector<uint64_t> batch_refs;  // list references per batch before deduplication
    uint32_t* hcol = nullptr;          // pinned host lists
    uint32_t* hcol_dev = nullptr;      // their device address
    uint64_t* rowptr = nullptr;        // HBM: offsets into hcol
    uint32_t* deg = nullptr;
    uint64_t* coff = nullptr;  // cache offset per vertex or kNotCached
    uint32_t* ccol = nullptr;  // HBM cache of lists
    uint64_t cached_vertices = 0, cache_entries = 0;
    uint32_t* stamp = nullptr;
    uint64_t* wrow = nullptr;
    uint32_t* wcol = nullptr;
    uint64_t wcap = 0;  // entries
    uint32_t* req = nullptr;
    uint32_t* nreq = nullptr;
    unsigned long long* bytes = nullptr;  // [host, cache]
    kz::Scratch cub;
    uint32_t epoch = 0;
    ~kz_paged_s() {
        if (hcol) (void)hipHostFree(hcol);
        for (void* p : {(void*)rowptr, (void*)deg, (void*)coff, (void*)ccol, (void*)stamp, (void*)wrow, (void*)wcol,
                        (void*)req, (void*)nreq, (void*)bytes, cub.ptr})
            if (p) (void)hipFree(p);
    }
};

using namespace kz;

extern "C" {

int kz_paged_create(const kz_csr* g, const kz_paged_cfg* cfg, kz_paged* out) {
    if (!g || !cfg || !out || !g->rowptr || (g->num_entries && !g->col)) KZ_FAIL(KZ_ERR_INVALID, "null argument");
    if (!g->oriented) KZ_FAIL(KZ_ERR_INVALID, "the paged form takes a degree-oriented graph (kz_orient_by_degree)");
    if (int e = ensure_ready()) return e;
    hipStream_t s = rt().stream;
    const uint32_t n = g->num_vertices;
    if (n >= (1u << 31)) KZ_FAIL(KZ_ERR_UNSUPPORTED, "the paged form supports < 2^31 vertices");
    const uint64_t m = g->num_entries;
    if (g->rowptr[n] != m) KZ_FAIL(KZ_ERR_INVALID, "rowptr[n] != num_entries");
    kz_paged p = new (std::nothrow) kz_paged_s();
    if (!p) KZ_FAIL(KZ_ERR_NOMEM, "paged graph");
    struct Guard {
        kz_paged& p;
        ~Guard() { delete p; }
    } guard{p};
    p->n = n;
    p->m = m;
    p->h_deg.resize(n);
    std::vector<uint32_t> indeg(n, 0);
#pragma omp parallel for schedule(static)
    for (int64_t v = 0; v < int64_t(n); ++v) {
        p->h_deg[v] = uint32_t(g->rowptr[v + 1] - g->rowptr[v]);
        for (uint64_t i = g->rowptr[v]; i < g->rowptr[v + 1]; ++i)
            __atomic_fetch_add(&indeg[g->col[i]], 1u, __ATOMIC_RELAXED);
    }
    // batches of consecutive roots whose undeduplicated list entries fit
    const uint64_t cap = cfg->list_bytes / 4;
    std::vector<uint64_t> cost(n);
#pragma omp parallel for schedule(dynamic, 4096)
    for (int64_t u = 0; u < int64_t(n); ++u) {
        uint64_t c = p->h_deg[u];
        for (uint64_t i = g->rowptr[u]; i < g->rowptr[u + 1]; ++i) c += p->h_deg[g->col[i]];
        cost[u] = c;
    }
    uint64_t acc = 0, refs = 0;
    for (uint32_t u = 0; u < n; ++u) {
        if (cost[u] > cap)
            KZ_FAIL(KZ_ERR_BUDGET, "root %u needs %llu list entries, more than list_bytes holds (%llu)", u,
                    (unsigned long long)cost[u], (unsigned long long)cap);
        if (acc + cost[u] > cap) {
            p->batch_end.push_back(u);
            p->batch_refs.push_back(refs);
            acc = refs = 0;
        }
        acc += cost[u];
        refs += 1 + p->h_deg[u];
    }
    if (n) {
        p->batch_end.push_back(n);
        p->batch_refs.push_back(refs);
    }
    // cache: total degree >= threshold, largest first, while it fits
    std::vector<uint32_t> cand;
    for (uint32_t v = 0; v < n; ++v)
        if (uint64_t(indeg[v]) + p->h_deg[v] >= cfg->cache_degree_threshold && p->h_deg[v]) cand.push_back(v);
    std::stable_sort(cand.begin(), cand.end(), [&](uint32_t a, uint32_t b) {
        return uint64_t(indeg[a]) + p->h_deg[a] > uint64_t(indeg[b]) + p->h_deg[b];
    });
    std::vector<uint64_t> coff(n, kNotCached);
    uint64_t centries = 0;
    std::vector<uint32_t> cached;
    for (uint32_t v : cand) {
        if (4 * (centries + p->h_deg[v]) > cfg->cache_bytes) break;
        coff[v] = centries;
        centries += p->h_deg[v];
        cached.push_back(v);
    }
    p->cached_vertices = cached.size();
    p->cache_entries = centries;
    // host lists (pinned, read by the kernels over PCIe)
    if (hipHostMalloc(reinterpret_cast<void**>(&p->hcol), col_alloc_bytes(m), hipHostMallocDefault) != hipSuccess) {
        (void)hipGetLastError();
        p->hcol = nullptr;
        KZ_FAIL(KZ_ERR_NOMEM, "pinned host lists (%zu bytes)", col_alloc_bytes(m));
    }
    if (hipHostGetDevicePointer(reinterpret_cast<void**>(&p->hcol_dev), p->hcol, 0) != hipSuccess) {
        (void)hipGetLastError();
        KZ_FAIL(KZ_ERR_DEVICE, "device address of the pinned host lists");
    }
    std::memcpy(p->hcol, g->col, 4 * m);
    std::memset(reinterpret_cast<char*>(p->hcol) + 4 * m, 0, col_alloc_bytes(m) - 4 * m);
    // HBM: offsets, degrees, cache, working arrays
    p->wcap = cap;
    hipError_t err = hipSuccess;
    auto H = [&](hipError_t e) {
        if (err == hipSuccess) err = e;
    };
    H(hipMalloc(&p->rowptr, 8 * (uint64_t(n) + 1)));
    H(hipMalloc(&p->deg, 4 * (uint64_t(n) + 1)));
    H(hipMalloc(&p->coff, 8 * (uint64_t(n) + 1)));
    H(hipMalloc(&p->ccol, col_alloc_bytes(centries)));
    H(hipMalloc(&p->stamp, 4 * (uint64_t(n) + 1)));
    H(hipMalloc(&p->wrow, 8 * (uint64_t(n) + 1)));
    H(hipMalloc(&p->wcol, col_alloc_bytes(cap)));
    H(hipMalloc(&p->req, 4 * (uint64_t(n) + 1)));
    H(hipMalloc(&p->nreq, 8));
    H(hipMalloc(&p->bytes, 16));
    if (err != hipSuccess) {
        (void)hipGetLastError();
        KZ_FAIL(KZ_ERR_NOMEM, "paged graph device arrays (%u vertices, %llu working entries, %llu cached)", n,
                (unsigned long long)cap, (unsigned long long)centries);
    }
    H(hipMemcpyAsync(p->rowptr, g->rowptr, 8 * (uint64_t(n) + 1), hipMemcpyHostToDevice, s));
    H(hipMemcpyAsync(p->deg, p->h_deg.data(), 4ull * n, hipMemcpyHostToDevice, s));
    H(hipMemcpyAsync(p->coff, coff.data(), 8ull * n, hipMemcpyHostToDevice, s));
    if (centries) {
        std::vector<uint32_t> cc(centries);
#pragma omp parallel for schedule(dynamic, 256)
        for (int64_t i = 0; i < int64_t(cached.size()); ++i) {
            const uint32_t v = cached[i];
            std::memcpy(cc.data() + coff[v], g->col + g->rowptr[v], 4ull * p->h_deg[v]);
        }
        H(hipMemcpyAsync(p->ccol, cc.data(), 4 * centries, hipMemcpyHostToDevice, s));
        H(hipStreamSynchronize(s));
    }
    if (n) hipLaunchKernelGGL(fill_u32, dim3(1024), dim3(256), 0, s, p->stamp, uint64_t(n) + 1, 0u);
    H(hipMemsetAsync(p->wcol, 0, col_alloc_bytes(cap), s));
    H(hipGetLastError());
    H(hipStreamSynchronize(s));
    if (err != hipSuccess) {
        (void)hipGetLastError();
        KZ_FAIL(KZ_ERR_DEVICE, "paged graph upload: %s", hipGetErrorString(err));
    }
    *out = p;
    p = nullptr;  // handed over
    return KZ_OK;
}

int kz_paged_count_cliques(kz_paged p, int k, uint64_t* count, kz_paged_stats* stats) {
    if (!p || !count) KZ_FAIL(KZ_ERR_INVALID, "null argument");
    if (k < 3 || k > 5) KZ_FAIL(KZ_ERR_UNSUPPORTED, "clique size %d outside 3..5", k);
    if (int e = ensure_ready()) return e;
    const auto t0 = std::chrono::steady_clock::now();
    hipStream_t s = rt().stream;
    kz_paged_stats st;
    std::memset(&st, 0, sizeof(st));
    st.cached_vertices = p->cached_vertices;
    uint64_t total = 0;
    const uint32_t n = p->n;
    const uint32_t grid = uint32_t(rt().num_cus) * 8;
    KZ_HIP(hipMemsetAsync(p->bytes, 0, 16, s));
    for (size_t b = 0; b < p->batch_end.size(); ++b) {
        const uint32_t r0 = b ? p->batch_end[b - 1] : 0u, r1 = p->batch_end[b];
        const uint32_t epoch = ++p->epoch;
        hipEvent_t e0, e1;
        KZ_HIP(hipEventCreate(&e0));
        KZ_HIP(hipEventCreate(&e1));
        KZ_HIP(hipEventRecord(e0, s));
        // requests, deduplicated by the batch stamp
        hipLaunchKernelGGL(request_kernel, dim3(grid), dim3(256), 0, s, r0, r1, p->rowptr, p->deg, p->hcol_dev, p->coff,
                           p->ccol, p->stamp, epoch);
        // working rows: scan of the stamped degrees (n + 1 entries)
        hipcub::CountingInputIterator<uint32_t> ids(0);
        hipcub::TransformInputIterator<uint64_t, StampedDeg, hipcub::CountingInputIterator<uint32_t>> wd(
            ids, StampedDeg{p->stamp, p->deg, epoch, n});
        size_t b1 = 0, b2 = 0;
        KZ_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, b1, wd, p->wrow, uint64_t(n) + 1, s));
        KZ_HIP(hipcub::DeviceSelect::If(nullptr, b2, ids, p->req, p->nreq, int(n), IsStamped{p->stamp, epoch}, s));
        if (int e = scratch_reserve(p->cub, std::max(b1, b2))) return e;
        KZ_HIP(hipcub::DeviceScan::ExclusiveSum(p->cub.ptr, b1, wd, p->wrow, uint64_t(n) + 1, s));
        KZ_HIP(hipcub::DeviceSelect::If(p->cub.ptr, b2, ids, p->req, p->nreq, int(n), IsStamped{p->stamp, epoch}, s));
        // the batch plan bounds the working entries by the buffer; checked
        // before any list is written
        uint32_t nr = 0;
        uint64_t went = 0;
        KZ_HIP(hipMemcpyAsync(&nr, p->nreq, 4, hipMemcpyDeviceToHost, s));
        KZ_HIP(hipMemcpyAsync(&went, p->wrow + n, 8, hipMemcpyDeviceToHost, s));
        KZ_HIP(hipStreamSynchronize(s));
        if (went > p->wcap)
            KZ_FAIL(KZ_ERR_STATE, "batch %zu needs %llu working entries (> %llu)", b, (unsigned long long)went,
                    (unsigned long long)p->wcap);
        hipLaunchKernelGGL(gather_kernel, dim3(grid), dim3(256), 0, s, p->req, p->nreq, p->rowptr, p->deg, p->hcol_dev,
                           p->coff, p->ccol, p->wrow, p->wcol, p->bytes);
        KZ_HIP(hipGetLastError());
        st.requests += nr;
        st.request_refs += p->batch_refs[b];
        // count the batch's roots on the working rows
        CliqueJob job;
        std::vector<CliqueQuery> q = {CliqueQuery{p->wrow, p->wcol, &p->h_deg, k, r0, r1}};
        if (int e = clique_job_prepare(job, q, default_buckets())) return e;
        uint64_t c = 0;
        double ms = 0;
        if (int e = clique_job_run(job, &c, &ms)) return e;
        KZ_HIP(hipEventRecord(e1, s));
        KZ_HIP(hipEventSynchronize(e1));
        float bms = 0;
        KZ_HIP(hipEventElapsedTime(&bms, e0, e1));
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        st.device_ms += bms;
        total += c;
        st.batches++;
    }
    unsigned long long by[2] = {0, 0};
    KZ_HIP(hipMemcpy(by, p->bytes, 16, hipMemcpyDeviceToHost));
    st.host_bytes = by[0];
    st.cache_bytes = by[1];
    st.wall_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    *count = total;
    if (stats) *stats = st;
    return KZ_OK;
}

int kz_paged_release(kz_paged p) {
    if (!p) return KZ_OK;
    if (rt().ready) (void)hipStreamSynchronize(rt().stream);
    delete p;
    return KZ_OK;
}

}  // extern "C"
