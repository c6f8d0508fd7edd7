// Device runtime of libkz_hip.so: device selection, the library stream,
// graph residency in HBM.
//
// Replaces, for one MI355X, the reference's process/NUMA runtime
// (src/distributed_sys.cc:12-41, src/shared_memory_sys.cc) and the
// per-socket CSR build of include/graph_loader.h:873-901: the whole CSR is
// copied once into HBM (288 GB) and stays resident; there is no partition,
// no remote fetch and no cache (src/engine.cc:89-647 have no counterpart).

#include <cstring>
#include <mutex>

#include "kz_device.h"

namespace kz {

static Runtime g_rt;
static std::mutex g_rt_mu;
Runtime& rt() { return g_rt; }

int ensure_ready() {
    if (!g_rt.ready) KZ_FAIL(KZ_ERR_STATE, "kz_init() has not been called");
    return KZ_OK;
}

int scratch_reserve(Scratch& s, size_t bytes) {
    if (bytes <= s.size) return KZ_OK;
    if (s.ptr) {
        KZ_HIP(hipStreamSynchronize(g_rt.stream));
        KZ_HIP(hipFree(s.ptr));
        s.ptr = nullptr;
        s.size = 0;
    }
    size_t want = bytes + bytes / 4;
    if (hipMalloc(&s.ptr, want) != hipSuccess) {
        (void)hipGetLastError();
        s.ptr = nullptr;
        if (hipMalloc(&s.ptr, bytes) != hipSuccess) {
            (void)hipGetLastError();
            s.ptr = nullptr;
            KZ_FAIL(KZ_ERR_NOMEM, "device scratch of %zu bytes", bytes);
        }
        want = bytes;
    }
    s.size = want;
    return KZ_OK;
}

}  // namespace kz

using namespace kz;

__global__ void kz_degree_kernel(const uint64_t* __restrict__ rowptr, uint32_t* __restrict__ deg, uint32_t n) {
    uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v < n) deg[v] = uint32_t(rowptr[v + 1] - rowptr[v]);
}

extern "C" {

int kz_init(int device) {
    std::lock_guard<std::mutex> lk(g_rt_mu);
    if (g_rt.ready) {
        if (g_rt.device == device) return KZ_OK;
        KZ_FAIL(KZ_ERR_STATE, "already initialised on device %d", g_rt.device);
    }
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) {
        (void)hipGetLastError();
        KZ_FAIL(KZ_ERR_DEVICE, "no HIP device available");
    }
    if (device < 0 || device >= count) KZ_FAIL(KZ_ERR_INVALID, "device %d out of range (%d devices)", device, count);
    KZ_HIP(hipSetDevice(device));
    hipDeviceProp_t p;
    KZ_HIP(hipGetDeviceProperties(&p, device));
    KZ_HIP(hipStreamCreateWithFlags(&g_rt.stream, hipStreamNonBlocking));
    g_rt.device = device;
    g_rt.num_cus = p.multiProcessorCount;
    g_rt.lds_per_block = p.sharedMemPerBlock;
    g_rt.ready = true;
    return KZ_OK;
}

int kz_shutdown(void) {
    std::lock_guard<std::mutex> lk(g_rt_mu);
    if (!g_rt.ready) return KZ_OK;
    KZ_HIP(hipStreamSynchronize(g_rt.stream));
    KZ_HIP(hipStreamDestroy(g_rt.stream));
    g_rt = Runtime();
    return KZ_OK;
}

int kz_get_device_info(kz_device_info* out) {
    if (!out) KZ_FAIL(KZ_ERR_INVALID, "null argument");
    if (int e = ensure_ready()) return e;
    hipDeviceProp_t p;
    KZ_HIP(hipGetDeviceProperties(&p, g_rt.device));
    std::memset(out, 0, sizeof(*out));
    std::strncpy(out->name, p.name, sizeof(out->name) - 1);
    std::strncpy(out->arch, p.gcnArchName, sizeof(out->arch) - 1);
    out->compute_units = p.multiProcessorCount;
    out->warp_size = p.warpSize;
    out->total_mem = p.totalGlobalMem;
    out->lds_per_block = p.sharedMemPerBlock;
    out->clock_khz = p.clockRate;
    out->mem_clock_khz = p.memoryClockRate;
    out->mem_bus_width = p.memoryBusWidth;
    out->l2_bytes = p.l2CacheSize;
    return KZ_OK;
}

int kz_graph_upload(const kz_csr* h, kz_graph* out) {
    if (!h || !out || !h->rowptr || (h->num_entries && !h->col)) KZ_FAIL(KZ_ERR_INVALID, "null argument");
    if (int e = ensure_ready()) return e;
    if (h->rowptr[h->num_vertices] != h->num_entries) KZ_FAIL(KZ_ERR_INVALID, "rowptr[n] != num_entries");
    kz_graph g = new (std::nothrow) kz_graph_s();
    if (!g) KZ_FAIL(KZ_ERR_NOMEM, "graph handle");
    DeviceGraph& d = g->g;
    d.n = h->num_vertices;
    d.nnz = h->num_entries;
    d.max_degree = h->max_degree;
    d.oriented = int(h->oriented);
    size_t rp_bytes = sizeof(uint64_t) * (size_t(d.n) + 1);
    size_t col_bytes = col_alloc_bytes(d.nnz);
    size_t deg_bytes = sizeof(uint32_t) * (d.n ? d.n : 1);
    auto fail = [&](const char* what) {
        hipFree(d.rowptr);
        hipFree(d.col);
        hipFree(d.deg);
        delete g;
        (void)hipGetLastError();
        kz::set_error("graph upload: %s", what);
        return KZ_ERR_NOMEM;
    };
    if (hipMalloc(&d.rowptr, rp_bytes) != hipSuccess) return fail("rowptr");
    if (hipMalloc(&d.col, col_bytes) != hipSuccess) return fail("col");
    if (hipMalloc(&d.deg, deg_bytes) != hipSuccess) return fail("deg");
    d.bytes = rp_bytes + col_bytes + deg_bytes;
    hipStream_t s = g_rt.stream;
    if (hipMemcpyAsync(d.rowptr, h->rowptr, rp_bytes, hipMemcpyHostToDevice, s) != hipSuccess ||
        (d.nnz && hipMemcpyAsync(d.col, h->col, sizeof(uint32_t) * d.nnz, hipMemcpyHostToDevice, s) != hipSuccess))
        return fail("copy");
    if (d.n) hipLaunchKernelGGL(kz_degree_kernel, dim3((d.n + 255) / 256), dim3(256), 0, s, d.rowptr, d.deg, d.n);
    if (hipStreamSynchronize(s) != hipSuccess) return fail("synchronize");
    d.h_deg.resize(d.n);
    uint32_t maxd = 0;
    for (uint32_t v = 0; v < d.n; ++v) {
        d.h_deg[v] = uint32_t(h->rowptr[v + 1] - h->rowptr[v]);
        maxd = d.h_deg[v] > maxd ? d.h_deg[v] : maxd;
    }
    d.max_degree = maxd;
    if (h->has_labels && h->labels && !d.oriented) {
        if (int e = fsm_set_labels(d, h->labels)) {
            kz_graph_release(g);
            return e;
        }
    }
    *out = g;
    return KZ_OK;
}

int kz_graph_release(kz_graph g) {
    if (!g) return KZ_OK;
    if (g_rt.ready) (void)hipStreamSynchronize(g_rt.stream);
    hipFree(g->g.rowptr);
    hipFree(g->g.col);
    hipFree(g->g.deg);
    if (g->g.fwd_rowptr && g->g.fwd_rowptr != g->g.rowptr) hipFree(g->g.fwd_rowptr);
    if (g->g.fwd_col && g->g.fwd_col != g->g.col) hipFree(g->g.fwd_col);
    delete g;
    return KZ_OK;
}

int kz_graph_info(kz_graph g, uint32_t* num_vertices, uint64_t* num_entries, uint32_t* max_degree, int* oriented) {
    if (!g) KZ_FAIL(KZ_ERR_INVALID, "null argument");
    if (num_vertices) *num_vertices = g->g.n;
    if (num_entries) *num_entries = g->g.nnz;
    if (max_degree) *max_degree = g->g.max_degree;
    if (oriented) *oriented = g->g.oriented;
    return KZ_OK;
}

int kz_graph_num_vertices(kz_graph g, uint32_t* out) {
    if (!g || !out) KZ_FAIL(KZ_ERR_INVALID, "null argument");
    *out = g->g.n;
    return KZ_OK;
}

}  // extern "C"
