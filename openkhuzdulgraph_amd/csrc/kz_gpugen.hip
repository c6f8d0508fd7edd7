// Graph generation and preprocessing on the GPU (kz_gpu_gen_*,
// kz_graph_download).
//
// Reference: the offline clean-up and orientation programs,
// datasets/live_journal/standardize.cc:66-114 (drop self-loops and duplicate
// edges, drop isolated vertices, relabel ascending) and
// toolkits/to_oriented_dataset_converter.cc:27-124 (rank vertices by degree,
// keep u -> v when rank(v) > rank(u), sort each list), which run on the CPU
// there.  Here the graphs of the host generators (kz_gen_*: the same
// counter-based draws, the same CSR bit for bit) are built in HBM:
//   1. every draw becomes the keys (u << 32 | v) and (v << 32 | u), self-loops
//      dropped; keys are scattered into 1024 buckets by their top bits (a
//      tile-local LDS histogram reserves each bucket's range with one atomic),
//      consecutive buckets are grouped into chunks below the radix sort's
//      item limit, and each chunk is radix-sorted and deduplicated (copies of
//      an edge share their bucket) and compacted in place;
//   2. row boundaries of the sorted keys give the degrees; a scan drops the
//      isolated vertices and relabels the rest ascending;
//   3. orientation: one radix sort of (degree << 32 | id) gives the ranks,
//      and the forward keys (rank u << 32 | rank v, rank v > rank u) go
//      through the same bucketed sort; rows come from the boundaries.

#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cstring>
#include <vector>

#include "kz_device.h"
#include "kz_kernels.h"

namespace kz {
namespace {

constexpr int kBucketBits = 10;
constexpr uint32_t kBuckets = 1u << kBucketBits;
uint64_t g_chunk_items = uint64_t(1) << 30;  // radix sort items per call (int)
constexpr uint32_t kTile = 4096;                     // items per scatter tile

int bits_for(uint64_t x) {  // smallest b with 2^b > x
    int b = 0;
    while (b < 64 && (uint64_t(1) << b) <= x) ++b;
    return b;
}

// ------------------------------------------------------------ key sources --
// A source yields up to one key per index j: bool operator()(j, &key).

struct ErSource {  // Erdos-Renyi G(n, m) draws (kz_gen_erdos_renyi)
    uint64_t key;
    uint32_t n;
    __device__ bool operator()(uint64_t j, uint64_t* out) const {
        const uint64_t r = draw64(key, j >> 1);
        const uint32_t u = uint32_t((uint64_t(uint32_t(r)) * n) >> 32);
        const uint32_t v = uint32_t((uint64_t(uint32_t(r >> 32)) * n) >> 32);
        if (u == v) return false;
        *out = (j & 1) ? (uint64_t(v) << 32 | u) : (uint64_t(u) << 32 | v);
        return true;
    }
};

struct RmatSource {  // RMAT draws (kz_gen_rmat)
    uint64_t key, ta, tb, tc;
    uint32_t scale, words;
    __device__ bool operator()(uint64_t j, uint64_t* out) const {
        const uint64_t i = j >> 1;
        uint32_t x = 0, y = 0;
        for (uint32_t w = 0; w < words; ++w) {
            const uint64_t r = draw64(key, i * words + w);
            for (int h = 0; h < 2; ++h) {
                const uint32_t level = 2 * w + uint32_t(h);
                if (level >= scale) break;
                const uint64_t q = h ? (r >> 32) : (r & 0xFFFFFFFFULL);
                const uint32_t bit = 1u << (scale - 1 - level);
                if (q < ta) {
                } else if (q < tb) {
                    y |= bit;
                } else if (q < tc) {
                    x |= bit;
                } else {
                    x |= bit;
                    y |= bit;
                }
            }
        }
        if (x == y) return false;
        *out = (j & 1) ? (uint64_t(y) << 32 | x) : (uint64_t(x) << 32 | y);
        return true;
    }
};

struct ChungLuSource {  // Chung-Lu draws (kz_gen_chung_lu) against the host's table
    uint64_t key;
    const double* cdf;  // n + 1 entries
    uint32_t n;
    double scale;       // cdf[n] / 2^32
    __device__ uint32_t pick(uint32_t r32) const {
        const double x = (double(r32) + 0.5) * scale;
        uint32_t lo = 0, hi = n;  // first k with cdf[k + 1] > x
        while (lo < hi) {
            const uint32_t mid = (lo + hi) >> 1;
            if (cdf[mid + 1] <= x) lo = mid + 1;
            else hi = mid;
        }
        return lo < n ? lo : n - 1;
    }
    __device__ bool operator()(uint64_t j, uint64_t* out) const {
        const uint64_t r = draw64(key, j >> 1);
        const uint32_t u = pick(uint32_t(r)), v = pick(uint32_t(r >> 32));
        if (u == v) return false;
        *out = (j & 1) ? (uint64_t(v) << 32 | u) : (uint64_t(u) << 32 | v);
        return true;
    }
};

struct CsrForwardSource {  // forward keys of a resident symmetric CSR
    const uint64_t* rowptr;
    const uint32_t* col;
    const uint32_t* rank;
    uint32_t n;
    __device__ bool operator()(uint64_t j, uint64_t* out) const {
        // row u of entry j: rowptr[lo] <= j < rowptr[hi] holds throughout
        uint32_t lo = 0, hi = n;
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) >> 1;
            if (rowptr[mid] <= j) lo = mid;
            else hi = mid;
        }
        const uint32_t ru = rank[lo], rv = rank[col[j]];
        if (rv <= ru) return false;
        *out = (uint64_t(ru) << 32) | rv;
        return true;
    }
};

struct ForwardSource {  // forward keys of the cleaned graph (orientation)
    const uint64_t* keys;  // sorted unique (u << 32 | v), original ids
    const uint32_t* newid;
    const uint32_t* rank;  // by new id
    __device__ bool operator()(uint64_t j, uint64_t* out) const {
        const uint64_t k = keys[j];
        const uint32_t ru = rank[newid[uint32_t(k >> 32)]], rv = rank[newid[uint32_t(k)]];
        if (rv <= ru) return false;
        *out = (uint64_t(ru) << 32) | rv;
        return true;
    }
};

// --------------------------------------------------------- bucketed sort --

template <class Src>
__global__ __launch_bounds__(256) void key_hist_kernel(Src src, uint64_t count, int shift,
                                                       unsigned long long* __restrict__ hist) {
    __shared__ uint32_t h[kBuckets];
    for (uint32_t i = threadIdx.x; i < kBuckets; i += blockDim.x) h[i] = 0;
    __syncthreads();
    for (uint64_t j = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; j < count;
         j += uint64_t(gridDim.x) * blockDim.x) {
        uint64_t k;
        if (src(j, &k)) atomicAdd(&h[uint32_t(k >> shift)], 1u);
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < kBuckets; i += blockDim.x)
        if (h[i]) atomicAdd(&hist[i], (unsigned long long)h[i]);
}

// Tile by tile: count the tile's keys per bucket in LDS, reserve each
// bucket's range with one global atomic, regenerate and place the keys.
template <class Src>
__global__ __launch_bounds__(256) void key_scatter_kernel(Src src, uint64_t count, int shift,
                                                          unsigned long long* __restrict__ cursor,
                                                          uint64_t* __restrict__ out) {
    __shared__ uint32_t h[kBuckets];
    __shared__ unsigned long long base[kBuckets];
    for (uint64_t t0 = uint64_t(blockIdx.x) * kTile; t0 < count; t0 += uint64_t(gridDim.x) * kTile) {
        for (uint32_t i = threadIdx.x; i < kBuckets; i += blockDim.x) h[i] = 0;
        __syncthreads();
        for (uint32_t t = threadIdx.x; t < kTile; t += blockDim.x) {
            uint64_t k;
            if (t0 + t < count && src(t0 + t, &k)) atomicAdd(&h[uint32_t(k >> shift)], 1u);
        }
        __syncthreads();
        for (uint32_t i = threadIdx.x; i < kBuckets; i += blockDim.x) {
            base[i] = h[i] ? atomicAdd(&cursor[i], (unsigned long long)h[i]) : 0ull;
            h[i] = 0;
        }
        __syncthreads();
        for (uint32_t t = threadIdx.x; t < kTile; t += blockDim.x) {
            uint64_t k;
            if (t0 + t < count && src(t0 + t, &k)) {
                const uint32_t b = uint32_t(k >> shift);
                out[base[b] + atomicAdd(&h[b], 1u)] = k;
            }
        }
        __syncthreads();
    }
}

__global__ void copy_u64_kernel(const uint64_t* __restrict__ in, uint64_t m, uint64_t* __restrict__ out) {
    for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < m; i += uint64_t(gridDim.x) * blockDim.x)
        out[i] = in[i];
}

struct Buffers {
    uint64_t* keys = nullptr;  // bucketed, then sorted (and compacted) in place
    uint64_t* tmp = nullptr;   // one chunk
    Scratch cub;
    unsigned long long* hist = nullptr;  // kBuckets
    unsigned long long* nsel = nullptr;  // unique count
    ~Buffers() {
        for (void* p : {(void*)keys, (void*)tmp, (void*)hist, (void*)nsel, cub.ptr})
            if (p) (void)hipFree(p);
    }
};

// Sorted keys of `src` over indices [0, count) (key bits [0, key_bits)) in
// b.keys[0, *total) (room for `capacity` keys); with dedupe, equal keys are
// kept once.
template <class Src>
int bucketed_sort(const Src& src, uint64_t count, int key_bits, bool dedupe, Buffers& b, uint64_t capacity,
                  uint64_t* total) {
    hipStream_t s = rt().stream;
    const int shift = std::max(0, key_bits - kBucketBits);
    const uint32_t grid = uint32_t(std::max(1, rt().num_cus * 4));
    KZ_HIP(hipMemsetAsync(b.hist, 0, sizeof(unsigned long long) * kBuckets, s));
    if (count) hipLaunchKernelGGL(key_hist_kernel<Src>, dim3(grid), dim3(256), 0, s, src, count, shift, b.hist);
    std::vector<unsigned long long> hist(kBuckets), off(kBuckets + 1, 0);
    KZ_HIP(hipMemcpyAsync(hist.data(), b.hist, sizeof(unsigned long long) * kBuckets, hipMemcpyDeviceToHost, s));
    KZ_HIP(hipStreamSynchronize(s));
    for (uint32_t i = 0; i < kBuckets; ++i) off[i + 1] = off[i] + hist[i];
    const uint64_t n = off[kBuckets];
    if (n > capacity) KZ_FAIL(KZ_ERR_INVALID, "%llu keys for a buffer of %llu", (unsigned long long)n,
                              (unsigned long long)capacity);
    for (uint32_t i = 0; i < kBuckets; ++i)
        if (hist[i] > g_chunk_items) KZ_FAIL(KZ_ERR_UNSUPPORTED, "one key bucket holds %llu keys (> %llu per sort)", hist[i],
                                    (unsigned long long)g_chunk_items);
    KZ_HIP(hipMemcpyAsync(b.hist, off.data(), sizeof(unsigned long long) * kBuckets, hipMemcpyHostToDevice, s));
    if (count) hipLaunchKernelGGL(key_scatter_kernel<Src>, dim3(grid), dim3(256), 0, s, src, count, shift, b.hist, b.keys);
    KZ_HIP(hipGetLastError());
    // chunks of consecutive buckets, each sorted into tmp and written back
    uint64_t write = 0;
    uint32_t c0 = 0;
    while (c0 < kBuckets) {
        uint32_t c1 = c0;
        while (c1 < kBuckets && off[c1 + 1] - off[c0] <= g_chunk_items) ++c1;
        const uint64_t lo = off[c0], len = off[c1] - off[c0];
        c0 = c1;
        if (!len) continue;
        size_t bytes = 0;
        KZ_HIP(hipcub::DeviceRadixSort::SortKeys(nullptr, bytes, b.keys + lo, b.tmp, int(len), 0, key_bits, s));
        size_t ubytes = 0;
        if (dedupe)
            KZ_HIP(hipcub::DeviceSelect::Unique(nullptr, ubytes, b.tmp, b.keys + write, b.nsel, int(len), s));
        if (int e = scratch_reserve(b.cub, std::max(bytes, ubytes))) return e;
        KZ_HIP(hipcub::DeviceRadixSort::SortKeys(b.cub.ptr, bytes, b.keys + lo, b.tmp, int(len), 0, key_bits, s));
        if (dedupe) {
            // keys[write, ...) lies below or inside this chunk's consumed input
            KZ_HIP(hipcub::DeviceSelect::Unique(b.cub.ptr, ubytes, b.tmp, b.keys + write, b.nsel, int(len), s));
            unsigned long long sel = 0;
            KZ_HIP(hipMemcpyAsync(&sel, b.nsel, 8, hipMemcpyDeviceToHost, s));
            KZ_HIP(hipStreamSynchronize(s));
            write += sel;
        } else {
            hipLaunchKernelGGL(copy_u64_kernel, dim3(1024), dim3(256), 0, s, b.tmp, len, b.keys + lo);
            write = lo + len;
        }
    }
    KZ_HIP(hipGetLastError());
    KZ_HIP(hipStreamSynchronize(s));
    *total = dedupe ? write : n;
    return KZ_OK;
}

// ---------------------------------------------------------------- rows --

// rows[u] = [start, end) of u's keys in the sorted array (0, 0 if absent)
__global__ void row_bounds_kernel(const uint64_t* __restrict__ keys, uint64_t m, uint64_t* __restrict__ start,
                                  uint64_t* __restrict__ end) {
    for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < m; i += uint64_t(gridDim.x) * blockDim.x) {
        const uint32_t u = uint32_t(keys[i] >> 32);
        if (i == 0 || uint32_t(keys[i - 1] >> 32) != u) start[u] = i;
        if (i + 1 == m || uint32_t(keys[i + 1] >> 32) != u) end[u] = i + 1;
    }
}

__global__ void degree_keep_kernel(const uint64_t* __restrict__ start, const uint64_t* __restrict__ end, uint32_t n,
                                   uint32_t* __restrict__ deg, uint32_t* __restrict__ keep) {
    for (uint32_t u = blockIdx.x * blockDim.x + threadIdx.x; u < n; u += gridDim.x * blockDim.x) {
        const uint32_t d = uint32_t(end[u] - start[u]);
        deg[u] = d;
        keep[u] = d ? 1u : 0u;
    }
}

__global__ void compact_degrees_kernel(const uint32_t* __restrict__ deg, const uint32_t* __restrict__ newid,
                                       uint32_t n, uint32_t* __restrict__ deg2) {
    for (uint32_t u = blockIdx.x * blockDim.x + threadIdx.x; u < n; u += gridDim.x * blockDim.x)
        if (deg[u]) deg2[newid[u]] = deg[u];
}

__global__ void relabel_col_kernel(const uint64_t* __restrict__ keys, uint64_t m, const uint32_t* __restrict__ newid,
                                   uint32_t* __restrict__ col) {
    for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < m; i += uint64_t(gridDim.x) * blockDim.x)
        col[i] = newid[uint32_t(keys[i])];
}

__global__ void low_word_col_kernel(const uint64_t* __restrict__ keys, uint64_t m, uint32_t* __restrict__ col) {
    for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < m; i += uint64_t(gridDim.x) * blockDim.x)
        col[i] = uint32_t(keys[i]);
}

__global__ void rank_key_kernel(const uint32_t* __restrict__ deg, uint32_t n, uint64_t* __restrict__ keys) {
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x)
        keys[v] = (uint64_t(deg[v]) << 32) | v;
}

__global__ void rank_from_sorted_kernel(const uint64_t* __restrict__ sorted, uint32_t n, uint32_t* __restrict__ rank) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        rank[uint32_t(sorted[i])] = i;
}

// Wave per vertex u: fdeg[rank u] = |{v in N(u) : rank v > rank u}|.
__global__ void rank_fdeg_kernel(const uint64_t* __restrict__ rowptr, const uint32_t* __restrict__ col,
                                 const uint32_t* __restrict__ rank, uint32_t n, uint32_t* __restrict__ fdeg) {
    const uint32_t u = (uint64_t(blockIdx.x) * blockDim.x + threadIdx.x) >> 6;
    const uint32_t lane = threadIdx.x & 63;
    if (u >= n) return;
    const uint64_t b = rowptr[u];
    const uint32_t d = uint32_t(rowptr[u + 1] - b);
    const uint32_t ru = rank[u];
    uint32_t c = 0;
    for (uint32_t k = lane; k < d; k += 64) c += rank[col[b + k]] > ru ? 1u : 0u;
    for (int off = 32; off; off >>= 1) c += __shfl_xor(c, off);
    if (lane == 0) fdeg[ru] = c;
}

__global__ void bounds_degree_kernel(const uint64_t* __restrict__ start, const uint64_t* __restrict__ end, uint32_t n,
                                     uint32_t* __restrict__ deg) {
    for (uint32_t u = blockIdx.x * blockDim.x + threadIdx.x; u < n; u += gridDim.x * blockDim.x)
        deg[u] = uint32_t(end[u] - start[u]);
}

struct U64Of {
    __host__ __device__ uint64_t operator()(uint32_t x) const { return x; }
};

int scan_u32_to_u64(const uint32_t* in, uint64_t* out, uint64_t n1, Scratch& tmp, hipStream_t s) {
    hipcub::TransformInputIterator<uint64_t, U64Of, const uint32_t*> it(in, U64Of());
    size_t bytes = 0;
    KZ_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, it, out, n1, s));
    if (int e = scratch_reserve(tmp, bytes)) return e;
    KZ_HIP(hipcub::DeviceScan::ExclusiveSum(tmp.ptr, bytes, it, out, n1, s));
    return KZ_OK;
}

int scan_u32(const uint32_t* in, uint32_t* out, uint64_t n1, Scratch& tmp, hipStream_t s) {
    size_t bytes = 0;
    KZ_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, in, out, n1, s));
    if (int e = scratch_reserve(tmp, bytes)) return e;
    KZ_HIP(hipcub::DeviceScan::ExclusiveSum(tmp.ptr, bytes, in, out, n1, s));
    return KZ_OK;
}

// Device allocation that is released unless `keep` is called.
template <class T>
struct DevArray {
    T* p = nullptr;
    ~DevArray() {
        if (p) (void)hipFree(p);
    }
    int alloc(size_t bytes, const char* what) {
        if (hipMalloc(&p, std::max<size_t>(bytes, 8)) != hipSuccess) {
            (void)hipGetLastError();
            p = nullptr;
            KZ_FAIL(KZ_ERR_NOMEM, "%s (%zu bytes)", what, bytes);
        }
        return KZ_OK;
    }
    T* keep() {
        T* r = p;
        p = nullptr;
        return r;
    }
};

// Builds the graph of draws [0, 2 * draws) of `src` over n_hint vertices.
template <class Src>
int build_graph(const Src& src, uint64_t draws, uint32_t n_hint, bool orient, kz_graph* out) {
    hipStream_t s = rt().stream;
    const uint64_t count = 2 * draws;
    const int kbits = 32 + bits_for(n_hint ? n_hint - 1 : 0);
    const uint32_t grid = uint32_t(std::max(1, rt().num_cus * 4));
    Buffers b;
    if (hipMalloc(&b.keys, 8 * std::max<uint64_t>(1, count)) != hipSuccess ||
        hipMalloc(&b.tmp, 8 * std::max<uint64_t>(1, std::min(count, g_chunk_items))) != hipSuccess ||
        hipMalloc(&b.hist, 8 * kBuckets) != hipSuccess || hipMalloc(&b.nsel, 8) != hipSuccess) {
        (void)hipGetLastError();
        KZ_FAIL(KZ_ERR_NOMEM, "generator keys for %llu draws", (unsigned long long)draws);
    }
    uint64_t m = 0;  // directed entries after dedupe
    if (int e = bucketed_sort(src, count, kbits, true, b, count, &m)) return e;
    // degrees, isolated vertices dropped, ascending relabel
    DevArray<uint64_t> start, end;
    DevArray<uint32_t> deg, keep, newid;
    if (int e = start.alloc(8ull * n_hint, "row starts")) return e;
    if (int e = end.alloc(8ull * n_hint, "row ends")) return e;
    if (int e = deg.alloc(4ull * n_hint, "degrees")) return e;
    if (int e = keep.alloc(4ull * (uint64_t(n_hint) + 1), "keep flags")) return e;
    if (int e = newid.alloc(4ull * (uint64_t(n_hint) + 1), "new ids")) return e;
    KZ_HIP(hipMemsetAsync(start.p, 0, 8ull * n_hint, s));
    KZ_HIP(hipMemsetAsync(end.p, 0, 8ull * n_hint, s));
    KZ_HIP(hipMemsetAsync(keep.p + n_hint, 0, 4, s));
    if (m) hipLaunchKernelGGL(row_bounds_kernel, dim3(grid), dim3(256), 0, s, b.keys, m, start.p, end.p);
    hipLaunchKernelGGL(degree_keep_kernel, dim3(grid), dim3(256), 0, s, start.p, end.p, n_hint, deg.p, keep.p);
    Scratch tmp;
    struct ScratchFree {
        Scratch& t;
        ~ScratchFree() {
            if (t.ptr) (void)hipFree(t.ptr);
        }
    } tmp_free{tmp};
    if (int e = scan_u32(keep.p, newid.p, uint64_t(n_hint) + 1, tmp, s)) return e;
    uint32_t n2 = 0;
    KZ_HIP(hipMemcpyAsync(&n2, newid.p + n_hint, 4, hipMemcpyDeviceToHost, s));
    KZ_HIP(hipStreamSynchronize(s));
    DevArray<uint32_t> deg2;
    if (int e = deg2.alloc(4ull * (uint64_t(n2) + 1), "compacted degrees")) return e;
    KZ_HIP(hipMemsetAsync(deg2.p + n2, 0, 4, s));
    hipLaunchKernelGGL(compact_degrees_kernel, dim3(grid), dim3(256), 0, s, deg.p, newid.p, n_hint, deg2.p);

    kz_graph g = new (std::nothrow) kz_graph_s();
    if (!g) KZ_FAIL(KZ_ERR_NOMEM, "graph handle");
    struct GraphGuard {
        kz_graph& g;
        ~GraphGuard() {
            if (g) kz_graph_release(g);
        }
    } guard{g};
    DeviceGraph& d = g->g;
    d.n = n2;
    DevArray<uint64_t> rowptr;
    DevArray<uint32_t> col, dout;
    if (!orient) {
        d.nnz = m;
        if (int e = rowptr.alloc(8ull * (uint64_t(n2) + 1), "rowptr")) return e;
        if (int e = col.alloc(col_alloc_bytes(m), "col")) return e;
        if (int e = scan_u32_to_u64(deg2.p, rowptr.p, uint64_t(n2) + 1, tmp, s)) return e;
        if (m) hipLaunchKernelGGL(relabel_col_kernel, dim3(grid), dim3(256), 0, s, b.keys, m, newid.p, col.p);
        d.deg = deg2.keep();
    } else {
        // rank by (degree, id), then the forward keys through the bucketed sort
        DevArray<uint64_t> rk, rs;
        DevArray<uint32_t> rank;
        if (int e = rk.alloc(8ull * n2, "rank keys")) return e;
        if (int e = rs.alloc(8ull * n2, "sorted rank keys")) return e;
        if (int e = rank.alloc(4ull * n2, "ranks")) return e;
        if (n2) {
            hipLaunchKernelGGL(rank_key_kernel, dim3(grid), dim3(256), 0, s, deg2.p, n2, rk.p);
            size_t bytes = 0;
            KZ_HIP(hipcub::DeviceRadixSort::SortKeys(nullptr, bytes, rk.p, rs.p, int(n2), 0, 64, s));
            if (int e = scratch_reserve(tmp, bytes)) return e;
            KZ_HIP(hipcub::DeviceRadixSort::SortKeys(tmp.ptr, bytes, rk.p, rs.p, int(n2), 0, 64, s));
            hipLaunchKernelGGL(rank_from_sorted_kernel, dim3(grid), dim3(256), 0, s, rs.p, n2, rank.p);
        }
        Buffers f;
        const uint64_t half = m / 2;
        if (hipMalloc(&f.keys, 8 * std::max<uint64_t>(1, half)) != hipSuccess ||
            hipMalloc(&f.tmp, 8 * std::max<uint64_t>(1, std::min(half, g_chunk_items))) != hipSuccess ||
            hipMalloc(&f.hist, 8 * kBuckets) != hipSuccess || hipMalloc(&f.nsel, 8) != hipSuccess) {
            (void)hipGetLastError();
            KZ_FAIL(KZ_ERR_NOMEM, "orientation keys for %llu edges", (unsigned long long)half);
        }
        const ForwardSource fsrc{b.keys, newid.p, rank.p};
        uint64_t mo = 0;
        if (int e = bucketed_sort(fsrc, m, 32 + bits_for(n2 ? n2 - 1 : 0), false, f, half, &mo)) return e;
        if (mo != half) KZ_FAIL(KZ_ERR_INVALID, "orientation kept %llu of %llu entries", (unsigned long long)mo,
                                (unsigned long long)m);
        d.nnz = mo;
        if (int e = rowptr.alloc(8ull * (uint64_t(n2) + 1), "rowptr")) return e;
        if (int e = col.alloc(col_alloc_bytes(mo), "col")) return e;
        if (int e = dout.alloc(4ull * (uint64_t(n2) + 1), "out-degrees")) return e;
        KZ_HIP(hipMemsetAsync(start.p, 0, 8ull * n2, s));
        KZ_HIP(hipMemsetAsync(end.p, 0, 8ull * n2, s));
        KZ_HIP(hipMemsetAsync(dout.p + n2, 0, 4, s));
        if (mo) hipLaunchKernelGGL(row_bounds_kernel, dim3(grid), dim3(256), 0, s, f.keys, mo, start.p, end.p);
        hipLaunchKernelGGL(bounds_degree_kernel, dim3(grid), dim3(256), 0, s, start.p, end.p, n2, dout.p);
        if (int e = scan_u32_to_u64(dout.p, rowptr.p, uint64_t(n2) + 1, tmp, s)) return e;
        if (mo) hipLaunchKernelGGL(low_word_col_kernel, dim3(grid), dim3(256), 0, s, f.keys, mo, col.p);
        KZ_HIP(hipStreamSynchronize(s));
        d.deg = dout.keep();
        d.oriented = 1;
    }
    KZ_HIP(hipGetLastError());
    d.h_deg.resize(n2);
    if (n2) KZ_HIP(hipMemcpyAsync(d.h_deg.data(), d.deg, 4ull * n2, hipMemcpyDeviceToHost, s));
    KZ_HIP(hipStreamSynchronize(s));
    d.rowptr = rowptr.keep();
    d.col = col.keep();
    uint32_t maxd = 0;
    for (uint32_t v = 0; v < n2; ++v) maxd = std::max(maxd, d.h_deg[v]);
    d.max_degree = maxd;
    d.bytes = 8ull * (uint64_t(n2) + 1) + col_alloc_bytes(d.nnz) + 4ull * std::max<uint32_t>(1, n2);
    *out = g;
    g = nullptr;  // handed over
    return KZ_OK;
}

}  // namespace

// Degree orientation of a resident symmetric graph
// (toolkits/to_oriented_dataset_converter.cc:27-124): vertices ranked by
// (degree, id), u -> v kept when rank v > rank u, ids replaced by ranks and
// every list sorted -- the CSR that kz_orient_by_degree builds on the host.
// The outputs are written only on success.
int orient_resident(const DeviceGraph& g, uint64_t** rowptr_out, uint32_t** col_out, uint32_t** deg_out,
                    std::vector<uint32_t>* h_deg, uint64_t* nnz_out) {
    hipStream_t s = rt().stream;
    const uint32_t n = g.n;
    if (n >= (1u << 31)) KZ_FAIL(KZ_ERR_UNSUPPORTED, "orientation supports < 2^31 vertices");
    const uint32_t grid = uint32_t(std::max(1, rt().num_cus * 4));
    Scratch tmp;
    struct ScratchFree {
        Scratch& t;
        ~ScratchFree() {
            if (t.ptr) (void)hipFree(t.ptr);
        }
    } tmp_free{tmp};
    DevArray<uint64_t> rk, rs, rowptr;
    DevArray<uint32_t> rank, fdeg, col;
    if (int e = rk.alloc(8ull * n, "rank keys")) return e;
    if (int e = rs.alloc(8ull * n, "sorted rank keys")) return e;
    if (int e = rank.alloc(4ull * n, "ranks")) return e;
    if (int e = fdeg.alloc(4ull * (uint64_t(n) + 1), "out-degrees")) return e;
    if (int e = rowptr.alloc(8ull * (uint64_t(n) + 1), "rowptr")) return e;
    KZ_HIP(hipMemsetAsync(fdeg.p + n, 0, 4, s));
    if (n) {
        hipLaunchKernelGGL(rank_key_kernel, dim3(grid), dim3(256), 0, s, g.deg, n, rk.p);
        size_t bytes = 0;
        KZ_HIP(hipcub::DeviceRadixSort::SortKeys(nullptr, bytes, rk.p, rs.p, int(n), 0, 64, s));
        if (int e = scratch_reserve(tmp, bytes)) return e;
        KZ_HIP(hipcub::DeviceRadixSort::SortKeys(tmp.ptr, bytes, rk.p, rs.p, int(n), 0, 64, s));
        hipLaunchKernelGGL(rank_from_sorted_kernel, dim3(grid), dim3(256), 0, s, rs.p, n, rank.p);
        hipLaunchKernelGGL(rank_fdeg_kernel, dim3((uint64_t(n) * 64 + 255) / 256), dim3(256), 0, s, g.rowptr, g.col,
                           rank.p, n, fdeg.p);
    }
    if (int e = scan_u32_to_u64(fdeg.p, rowptr.p, uint64_t(n) + 1, tmp, s)) return e;
    uint64_t fnnz = 0;
    KZ_HIP(hipMemcpyAsync(&fnnz, rowptr.p + n, 8, hipMemcpyDeviceToHost, s));
    KZ_HIP(hipStreamSynchronize(s));
    if (int e = col.alloc(col_alloc_bytes(fnnz), "forward lists")) return e;
    if (fnnz) {
        Buffers f;
        if (hipMalloc(&f.keys, 8 * fnnz) != hipSuccess ||
            hipMalloc(&f.tmp, 8 * std::min(fnnz, g_chunk_items)) != hipSuccess ||
            hipMalloc(&f.hist, 8 * kBuckets) != hipSuccess || hipMalloc(&f.nsel, 8) != hipSuccess) {
            (void)hipGetLastError();
            KZ_FAIL(KZ_ERR_NOMEM, "orientation keys for %llu edges", (unsigned long long)fnnz);
        }
        const CsrForwardSource src{g.rowptr, g.col, rank.p, n};
        uint64_t mo = 0;
        if (int e = bucketed_sort(src, g.nnz, 32 + bits_for(n ? n - 1 : 0), false, f, fnnz, &mo)) return e;
        if (mo != fnnz) KZ_FAIL(KZ_ERR_INVALID, "orientation kept %llu of %llu forward entries",
                                (unsigned long long)mo, (unsigned long long)fnnz);
        hipLaunchKernelGGL(low_word_col_kernel, dim3(grid), dim3(256), 0, s, f.keys, fnnz, col.p);
        KZ_HIP(hipGetLastError());
        KZ_HIP(hipStreamSynchronize(s));
    }
    std::vector<uint32_t> hd(n);
    if (n) KZ_HIP(hipMemcpyAsync(hd.data(), fdeg.p, 4ull * n, hipMemcpyDeviceToHost, s));
    KZ_HIP(hipStreamSynchronize(s));
    *rowptr_out = rowptr.keep();
    *col_out = col.keep();
    if (deg_out) *deg_out = fdeg.keep();
    h_deg->swap(hd);
    *nnz_out = fnnz;
    return KZ_OK;
}

}  // namespace kz

using namespace kz;

extern "C" {

int kz_gpu_gen_erdos_renyi(uint32_t n, uint64_t m, uint64_t seed, int orient, kz_graph* out) {
    if (!out || n < 2) KZ_FAIL(KZ_ERR_INVALID, "erdos_renyi needs n >= 2");
    if (int e = ensure_ready()) return e;
    return build_graph(ErSource{stream_key(seed, 0x4552ULL), n}, m, n, orient != 0, out);
}

int kz_gpu_gen_rmat(uint32_t scale, uint32_t edge_factor, double a, double b, double c, uint64_t seed, int orient,
                    kz_graph* out) {
    if (!out || scale < 1 || scale > 31 || a < 0 || b < 0 || c < 0 || a + b + c > 1.0)
        KZ_FAIL(KZ_ERR_INVALID, "rmat parameters out of range");
    if (int e = ensure_ready()) return e;
    RmatSource src;
    src.key = stream_key(seed, 0x524D4154ULL);
    src.ta = uint64_t(a * 4294967296.0);
    src.tb = uint64_t((a + b) * 4294967296.0);
    src.tc = uint64_t((a + b + c) * 4294967296.0);
    src.scale = scale;
    src.words = (scale + 1) / 2;
    return build_graph(src, uint64_t(edge_factor) << scale, uint32_t(1u << scale), orient != 0, out);
}

int kz_gpu_gen_chung_lu(uint32_t n, uint64_t num_draws, double exponent, uint32_t max_weight, uint64_t seed,
                        int orient, kz_graph* out) {
    if (!out) KZ_FAIL(KZ_ERR_INVALID, "null argument");
    if (int e = ensure_ready()) return e;
    std::vector<double> cdf;
    if (int e = chung_lu_cdf(n, num_draws, exponent, max_weight, &cdf)) return e;
    DevArray<double> dcdf;
    if (int e = dcdf.alloc(8 * cdf.size(), "chung-lu table")) return e;
    KZ_HIP(hipMemcpy(dcdf.p, cdf.data(), 8 * cdf.size(), hipMemcpyHostToDevice));
    const ChungLuSource src{stream_key(seed, 0x434C55ULL), dcdf.p, n, cdf[n] / 4294967296.0};
    return build_graph(src, num_draws, n, orient != 0, out);
}

int kz_set_sort_chunk(uint64_t items) {
    if (items < 1024 || items > (uint64_t(1) << 30)) KZ_FAIL(KZ_ERR_INVALID, "sort chunk must be in [2^10, 2^30]");
    g_chunk_items = items;
    return KZ_OK;
}

int kz_graph_orient(kz_graph sym, kz_graph* out) {
    if (!sym || !out) KZ_FAIL(KZ_ERR_INVALID, "null argument");
    if (sym->g.oriented) KZ_FAIL(KZ_ERR_INVALID, "the graph is already oriented");
    if (int e = ensure_ready()) return e;
    kz_graph g = new (std::nothrow) kz_graph_s();
    if (!g) KZ_FAIL(KZ_ERR_NOMEM, "graph handle");
    DeviceGraph& d = g->g;
    if (int e = orient_resident(sym->g, &d.rowptr, &d.col, &d.deg, &d.h_deg, &d.nnz)) {
        delete g;
        return e;
    }
    d.n = sym->g.n;
    d.oriented = 1;
    for (uint32_t x : d.h_deg) d.max_degree = std::max(d.max_degree, x);
    d.bytes = 8ull * (uint64_t(d.n) + 1) + col_alloc_bytes(d.nnz) + 4ull * (uint64_t(d.n) + 1);
    *out = g;
    return KZ_OK;
}

int kz_graph_download(kz_graph g, kz_csr* out) {
    if (!g || !out) KZ_FAIL(KZ_ERR_INVALID, "null argument");
    const DeviceGraph& d = g->g;
    if (!csr_alloc(out, d.n, d.nnz, false)) KZ_FAIL(KZ_ERR_NOMEM, "host csr");
    if (hipMemcpy(out->rowptr, d.rowptr, 8ull * (uint64_t(d.n) + 1), hipMemcpyDeviceToHost) != hipSuccess ||
        (d.nnz && hipMemcpy(out->col, d.col, 4ull * d.nnz, hipMemcpyDeviceToHost) != hipSuccess)) {
        (void)hipGetLastError();
        kz_csr_release(out);
        KZ_FAIL(KZ_ERR_DEVICE, "graph download");
    }
    out->max_degree = d.max_degree;
    out->oriented = uint32_t(d.oriented);
    return KZ_OK;
}

}  // extern "C"
