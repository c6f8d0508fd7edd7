// Host-side graph layer of libkz_hip.so: seeded generators, clean-up into a
// symmetric CSR, degree orientation, .dgraph read/write and structural checks.
//
// Reference behaviour followed (paths relative to the OpenKhuzdulGraph tree):
//   * clean-up: datasets/live_journal/standardize.cc:66-114 (drop self-loops
//     and duplicates, compact ids in ascending original order) and
//     toolkits/converter.cc:237-263 (store both directions, sorted by
//     (src, dst));
//   * orientation: toolkits/to_oriented_dataset_converter.cc:27-124;
//   * .dgraph layout: include/graph_loader.h:31-42 and the loader at
//     graph_loader.h:706-961 (degree sum must equal num_edges, :891-898);
//   * checks: toolkits/check_graph.cc:44-68.
// The code is written for multi-core hosts with OpenMP: every pass is a
// parallel loop over vertices or edge draws, with atomics only for degree
// counting and bucket placement; per-list sorting makes the output
// independent of thread interleaving.

#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>
#include <omp.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "kz_internal.h"

namespace kz {

static thread_local std::string g_error;

void set_error(const char* fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_error = buf;
}
const char* get_error() { return g_error.c_str(); }

bool csr_alloc(kz_csr* g, uint32_t n, uint64_t entries, bool labels) {
    std::memset(g, 0, sizeof(*g));
    g->num_vertices = n;
    g->num_entries = entries;
    g->rowptr = static_cast<uint64_t*>(std::malloc(sizeof(uint64_t) * (size_t(n) + 1)));
    g->col = static_cast<uint32_t*>(std::malloc(sizeof(uint32_t) * (entries ? entries : 1)));
    if (labels) {
        g->labels = static_cast<uint32_t*>(std::malloc(sizeof(uint32_t) * (n ? n : 1)));
        g->has_labels = 1;
    }
    if (!g->rowptr || !g->col || (labels && !g->labels)) {
        std::free(g->rowptr);
        std::free(g->col);
        std::free(g->labels);
        std::memset(g, 0, sizeof(*g));
        return false;
    }
    return true;
}

// Exclusive prefix sum of counts[0..n) into out[0..n], returns the total.
template <typename CountT>
static uint64_t exclusive_scan(const CountT* counts, uint64_t* out, size_t n) {
    int nt = omp_get_max_threads();
    std::vector<uint64_t> part(size_t(nt) + 1, 0);
#pragma omp parallel num_threads(nt)
    {
        int t = omp_get_thread_num();
        int T = omp_get_num_threads();
        size_t lo = n * t / T, hi = n * (t + 1) / T;
        uint64_t s = 0;
        for (size_t i = lo; i < hi; ++i) s += counts[i];
        part[t + 1] = s;
#pragma omp barrier
#pragma omp single
        for (int i = 0; i < T; ++i) part[i + 1] += part[i];
        uint64_t run = part[t];
        for (size_t i = lo; i < hi; ++i) {
            out[i] = run;
            run += counts[i];
        }
    }
    out[n] = part[nt];
    return part[nt];
}

// Generic clean-up: `edge(i, &u, &v)` yields undirected draw i (i < m).
// Both passes regenerate the draws instead of storing them.
template <typename EdgeFn>
static int clean_symmetric(uint64_t m, uint32_t n_hint, EdgeFn edge, kz_csr* out) {
    std::vector<uint32_t> deg(n_hint, 0);
    bool bad = false;
#pragma omp parallel for schedule(static) reduction(|| : bad)
    for (uint64_t i = 0; i < m; ++i) {
        uint32_t u, v;
        edge(i, &u, &v);
        if (u >= n_hint || v >= n_hint) {
            bad = true;
            continue;
        }
        if (u == v) continue;  // self-loops are dropped (standardize.cc:99)
        __atomic_fetch_add(&deg[u], 1u, __ATOMIC_RELAXED);
        __atomic_fetch_add(&deg[v], 1u, __ATOMIC_RELAXED);
    }
    if (bad) KZ_FAIL(KZ_ERR_INVALID, "edge endpoint >= num_vertices_hint (%u)", n_hint);

    std::vector<uint64_t> off(size_t(n_hint) + 1);
    uint64_t total = exclusive_scan(deg.data(), off.data(), n_hint);
    uint32_t* tmp = static_cast<uint32_t*>(std::malloc(sizeof(uint32_t) * (total ? total : 1)));
    if (!tmp) KZ_FAIL(KZ_ERR_NOMEM, "clean-up buffer of %llu entries", (unsigned long long)total);
    std::vector<uint64_t> pos(off.begin(), off.end() - 1);
#pragma omp parallel for schedule(static)
    for (uint64_t i = 0; i < m; ++i) {
        uint32_t u, v;
        edge(i, &u, &v);
        if (u == v) continue;
        tmp[__atomic_fetch_add(&pos[u], 1ull, __ATOMIC_RELAXED)] = v;
        tmp[__atomic_fetch_add(&pos[v], 1ull, __ATOMIC_RELAXED)] = u;
    }
    // Sort + dedupe each list in place; record the unique degree.
#pragma omp parallel for schedule(dynamic, 1024)
    for (uint64_t u = 0; u < n_hint; ++u) {
        uint32_t* b = tmp + off[u];
        uint32_t* e = tmp + off[u + 1];
        std::sort(b, e);
        deg[u] = uint32_t(std::unique(b, e) - b);
    }
    // Drop isolated vertices; relabel ascending (standardize.cc:74-86).
    std::vector<uint32_t> keep(n_hint);
#pragma omp parallel for schedule(static)
    for (uint64_t u = 0; u < n_hint; ++u) keep[u] = deg[u] ? 1u : 0u;
    std::vector<uint64_t> newid(size_t(n_hint) + 1);
    uint32_t n2 = uint32_t(exclusive_scan(keep.data(), newid.data(), n_hint));
    std::vector<uint32_t> deg2(n2);
#pragma omp parallel for schedule(static)
    for (uint64_t u = 0; u < n_hint; ++u)
        if (deg[u]) deg2[newid[u]] = deg[u];
    if (!csr_alloc(out, n2, 0, false)) {
        std::free(tmp);
        KZ_FAIL(KZ_ERR_NOMEM, "csr rowptr");
    }
    uint64_t entries = exclusive_scan(deg2.data(), out->rowptr, n2);
    std::free(out->col);
    out->col = static_cast<uint32_t*>(std::malloc(sizeof(uint32_t) * (entries ? entries : 1)));
    if (!out->col) {
        std::free(tmp);
        kz_csr_release(out);
        KZ_FAIL(KZ_ERR_NOMEM, "csr col of %llu entries", (unsigned long long)entries);
    }
    out->num_entries = entries;
    uint32_t maxd = 0;
#pragma omp parallel for schedule(dynamic, 1024) reduction(max : maxd)
    for (uint64_t u = 0; u < n_hint; ++u) {
        uint32_t d = deg[u];
        if (!d) continue;
        const uint32_t* s = tmp + off[u];
        uint32_t* dst = out->col + out->rowptr[newid[u]];
        for (uint32_t k = 0; k < d; ++k) dst[k] = uint32_t(newid[s[k]]);
        maxd = std::max(maxd, d);
    }
    out->max_degree = maxd;
    out->oriented = 0;
    std::free(tmp);
    return KZ_OK;
}

}  // namespace kz

using namespace kz;

extern "C" {

const char* kz_last_error(void) { return kz::get_error(); }

void kz_csr_release(kz_csr* g) {
    if (!g) return;
    std::free(g->rowptr);
    std::free(g->col);
    std::free(g->labels);
    std::memset(g, 0, sizeof(*g));
}

int kz_csr_from_edges(const uint32_t* src, const uint32_t* dst, uint64_t num_edges,
                      uint32_t num_vertices_hint, kz_csr* out) {
    if (!out || (num_edges && (!src || !dst))) KZ_FAIL(KZ_ERR_INVALID, "null argument");
    auto fn = [&](uint64_t i, uint32_t* u, uint32_t* v) {
        *u = src[i];
        *v = dst[i];
    };
    return clean_symmetric(num_edges, num_vertices_hint, fn, out);
}

int kz_gen_erdos_renyi(uint32_t n, uint64_t m, uint64_t seed, kz_csr* out) {
    if (!out || n < 2) KZ_FAIL(KZ_ERR_INVALID, "erdos_renyi needs n >= 2");
    const uint64_t key = stream_key(seed, 0x4552ULL);
    auto fn = [&](uint64_t i, uint32_t* u, uint32_t* v) {
        uint64_t r = draw64(key, i);
        // Lemire multiply-shift maps a 32-bit draw onto [0, n) without bias
        // worth noting at these sizes and exactly when n is a power of two.
        *u = uint32_t((uint64_t(uint32_t(r)) * n) >> 32);
        *v = uint32_t((uint64_t(uint32_t(r >> 32)) * n) >> 32);
    };
    return clean_symmetric(m, n, fn, out);
}

int kz_gen_rmat(uint32_t scale, uint32_t edge_factor, double a, double b, double c,
                uint64_t seed, kz_csr* out) {
    if (!out || scale < 1 || scale > 31 || a < 0 || b < 0 || c < 0 || a + b + c > 1.0)
        KZ_FAIL(KZ_ERR_INVALID, "rmat parameters out of range");
    const uint64_t key = stream_key(seed, 0x524D4154ULL);
    const uint64_t ta = uint64_t(a * 4294967296.0);
    const uint64_t tb = uint64_t((a + b) * 4294967296.0);
    const uint64_t tc = uint64_t((a + b + c) * 4294967296.0);
    const uint32_t words = (scale + 1) / 2;  // two 32-bit quadrant draws per word
    const uint64_t m = uint64_t(edge_factor) << scale;
    auto fn = [&](uint64_t i, uint32_t* u, uint32_t* v) {
        uint32_t x = 0, y = 0;
        for (uint32_t w = 0; w < words; ++w) {
            uint64_t r = draw64(key, i * words + w);
            for (int h = 0; h < 2; ++h) {
                uint32_t level = 2 * w + h;
                if (level >= scale) break;
                uint64_t q = h ? (r >> 32) : (r & 0xFFFFFFFFULL);
                uint32_t bit = 1u << (scale - 1 - level);
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
        *u = x;
        *v = y;
    };
    return clean_symmetric(m, uint32_t(1u << scale), fn, out);
}

int kz_gen_chung_lu(uint32_t n, uint64_t num_draws, double exponent, uint32_t max_weight,
                    uint64_t seed, kz_csr* out) {
    if (!out) KZ_FAIL(KZ_ERR_INVALID, "null argument");
    std::vector<double> cdf;
    if (int e = chung_lu_cdf(n, num_draws, exponent, max_weight, &cdf)) return e;
    const double total = cdf[n];
    const uint64_t key = stream_key(seed, 0x434C55ULL);
    auto pick = [&](uint32_t r32) {
        double x = (double(r32) + 0.5) * (total / 4294967296.0);
        uint64_t k = uint64_t(std::upper_bound(cdf.begin() + 1, cdf.end(), x) - (cdf.begin() + 1));
        return uint32_t(k < n ? k : n - 1);
    };
    auto fn = [&](uint64_t i, uint32_t* u, uint32_t* v) {
        uint64_t r = draw64(key, i);
        *u = pick(uint32_t(r));
        *v = pick(uint32_t(r >> 32));
    };
    return clean_symmetric(num_draws, n, fn, out);
}

}  // extern "C"

int kz::chung_lu_cdf(uint32_t n, uint64_t num_draws, double exponent, uint32_t max_weight,
                     std::vector<double>* cdf_out) {
    if (n < 2 || exponent <= 1.0 || max_weight < 1) KZ_FAIL(KZ_ERR_INVALID, "chung_lu parameters out of range");
    // Weights w_i = max_weight * (i0 / (i + i0))^(1/(exponent-1)); i0 is
    // chosen so that sum(w) = 2 * num_draws (expected degree sum).
    const double alpha = 1.0 / (exponent - 1.0);
    const double target = 2.0 * double(num_draws);
    auto weight_sum = [&](double i0) {
        double s = 0;
#pragma omp parallel for reduction(+ : s)
        for (uint64_t i = 0; i < n; ++i) s += double(max_weight) * std::pow(i0 / (double(i) + i0), alpha);
        return s;
    };
    double lo = 1e-9, hi = double(n) * 1e6;
    if (weight_sum(hi) < target) KZ_FAIL(KZ_ERR_INVALID, "num_draws too large for n and max_weight");
    for (int it = 0; it < 200; ++it) {
        double mid = std::sqrt(lo * hi);
        (weight_sum(mid) < target ? lo : hi) = mid;
    }
    const double i0 = hi;
    std::vector<double>& cdf = *cdf_out;
    cdf.assign(size_t(n) + 1, 0.0);
    for (uint64_t i = 0; i < n; ++i)
        cdf[i + 1] = cdf[i] + double(max_weight) * std::pow(i0 / (double(i) + i0), alpha);
    return KZ_OK;
}

extern "C" {

int kz_orient_by_degree(const kz_csr* sym, kz_csr* out) {
    if (!sym || !out || !sym->rowptr) KZ_FAIL(KZ_ERR_INVALID, "null graph");
    const uint32_t n = sym->num_vertices;
    if (sym->num_entries % 2) KZ_FAIL(KZ_ERR_INVALID, "symmetric graph must have an even entry count");
    // Rank by (degree asc, id asc): a stable counting sort by degree
    // (to_oriented_dataset_converter.cc:47-67).
    std::vector<uint32_t> rank(n);
    {
        uint32_t maxd = 0;
        for (uint32_t u = 0; u < n; ++u) maxd = std::max<uint32_t>(maxd, uint32_t(sym->rowptr[u + 1] - sym->rowptr[u]));
        std::vector<uint64_t> bucket(size_t(maxd) + 2, 0);
        for (uint32_t u = 0; u < n; ++u) bucket[sym->rowptr[u + 1] - sym->rowptr[u] + 1]++;
        for (size_t d = 1; d < bucket.size(); ++d) bucket[d] += bucket[d - 1];
        for (uint32_t u = 0; u < n; ++u) rank[u] = uint32_t(bucket[sym->rowptr[u + 1] - sym->rowptr[u]]++);
    }
    std::vector<uint32_t> outdeg(n);
#pragma omp parallel for schedule(dynamic, 1024)
    for (uint64_t u = 0; u < n; ++u) {
        uint32_t c = 0, r = rank[u];
        for (uint64_t e = sym->rowptr[u]; e < sym->rowptr[u + 1]; ++e) c += rank[sym->col[e]] > r;
        outdeg[r] = c;
    }
    if (!csr_alloc(out, n, sym->num_entries / 2, false)) KZ_FAIL(KZ_ERR_NOMEM, "oriented csr");
    uint64_t total = exclusive_scan(outdeg.data(), out->rowptr, n);
    if (total != sym->num_entries / 2) {
        kz_csr_release(out);
        KZ_FAIL(KZ_ERR_INVALID, "graph is not symmetric (%llu oriented of %llu entries)",
                (unsigned long long)total, (unsigned long long)sym->num_entries);
    }
    uint32_t maxd = 0;
#pragma omp parallel for schedule(dynamic, 1024) reduction(max : maxd)
    for (uint64_t u = 0; u < n; ++u) {
        uint32_t r = rank[u];
        uint32_t* dst = out->col + out->rowptr[r];
        uint32_t k = 0;
        for (uint64_t e = sym->rowptr[u]; e < sym->rowptr[u + 1]; ++e) {
            uint32_t rv = rank[sym->col[e]];
            if (rv > r) dst[k++] = rv;
        }
        std::sort(dst, dst + k);  // to_oriented_dataset_converter.cc:98-104
        maxd = std::max(maxd, k);
    }
    out->max_degree = maxd;
    out->oriented = 1;
    return KZ_OK;
}

int kz_csr_check(const kz_csr* g) {
    if (!g || !g->rowptr || (g->num_entries && !g->col)) KZ_FAIL(KZ_ERR_INVALID, "null graph");
    const uint32_t n = g->num_vertices;
    if (g->rowptr[0] != 0 || g->rowptr[n] != g->num_entries)
        KZ_FAIL(KZ_ERR_INVALID, "degree sum %llu != entries %llu", (unsigned long long)g->rowptr[n],
                (unsigned long long)g->num_entries);
    int err = 0;
    uint64_t err_v = 0;
#pragma omp parallel for schedule(dynamic, 1024)
    for (uint64_t u = 0; u < n; ++u) {
        uint64_t b = g->rowptr[u], e = g->rowptr[u + 1];
        if (e < b) {
            __atomic_store_n(&err, 1, __ATOMIC_RELAXED);
            err_v = u;
            continue;
        }
        for (uint64_t k = b; k < e; ++k) {
            uint32_t v = g->col[k];
            int code = 0;
            if (v >= n) code = 2;
            else if (v == u) code = 3;  // check_graph.cc:57
            else if (k > b && g->col[k - 1] >= v) code = 4;  // check_graph.cc:54
            else if (g->oriented && v < u) code = 5;
            else if (!g->oriented) {
                // reverse edge must exist (check_graph.cc:62-65)
                const uint32_t* rb = g->col + g->rowptr[v];
                const uint32_t* re = g->col + g->rowptr[v + 1];
                if (!std::binary_search(rb, re, uint32_t(u))) code = 6;
            }
            if (code) {
                __atomic_store_n(&err, code, __ATOMIC_RELAXED);
                err_v = u;
                break;
            }
        }
    }
    static const char* what[] = {"", "negative degree", "id out of range", "self-loop",
                                 "list not strictly increasing", "oriented edge with dst < src",
                                 "missing reverse edge"};
    if (err) KZ_FAIL(KZ_ERR_INVALID, "%s at vertex %llu", what[err], (unsigned long long)err_v);
    return KZ_OK;
}

int kz_csr_set_labels(kz_csr* g, const uint32_t* labels) {
    if (!g || !g->rowptr || (!labels && g->num_vertices)) KZ_FAIL(KZ_ERR_INVALID, "null argument");
    for (uint32_t v = 0; v < g->num_vertices; ++v)
        if (labels[v] >= KZ_MAX_LABELS)
            KZ_FAIL(KZ_ERR_INVALID, "label %u of vertex %u is not below %d", labels[v], v, KZ_MAX_LABELS);
    uint32_t* copy = static_cast<uint32_t*>(std::malloc(sizeof(uint32_t) * (g->num_vertices ? g->num_vertices : 1)));
    if (!copy) KZ_FAIL(KZ_ERR_NOMEM, "labels");
    if (g->num_vertices) std::memcpy(copy, labels, sizeof(uint32_t) * g->num_vertices);
    std::free(g->labels);
    g->labels = copy;
    g->has_labels = 1;
    return KZ_OK;
}

uint64_t kz_csr_digest(const kz_csr* g) {
    if (!g || !g->rowptr) return 0;
    // Word-wise mix over (n, entries, rowptr[], col[]); chunks are hashed in
    // parallel and folded in order, so the value is thread-count independent.
    auto fold = [](uint64_t h, uint64_t w) { return mix64(h ^ (w + kGolden)); };
    uint64_t h = fold(fold(0x6B7A6469676573ULL, g->num_vertices), g->num_entries);
    const size_t chunk = size_t(1) << 20;
    auto hash_words = [&](const uint64_t* p, size_t nw, uint64_t h0) {
        size_t nch = (nw + chunk - 1) / chunk;
        std::vector<uint64_t> part(nch);
#pragma omp parallel for schedule(static)
        for (size_t c = 0; c < nch; ++c) {
            uint64_t x = c + 1;
            size_t e = std::min(nw, (c + 1) * chunk);
            for (size_t i = c * chunk; i < e; ++i) x = fold(x, p[i]);
            part[c] = x;
        }
        for (size_t c = 0; c < nch; ++c) h0 = fold(h0, part[c]);
        return h0;
    };
    h = hash_words(g->rowptr, size_t(g->num_vertices) + 1, h);
    size_t ncol = size_t(g->num_entries);
    h = hash_words(reinterpret_cast<const uint64_t*>(g->col), ncol / 2, h);
    if (ncol & 1) h = fold(h, g->col[ncol - 1]);
    return h;
}

// ---------------------------------------------------------------- .dgraph --

#pragma pack(push, 1)
struct DGraphHeader {  // graph_loader.h:37-42
    uint32_t num_vertices;
    uint64_t num_edges;
    int32_t has_node_label;
    int32_t has_edge_label;
};
#pragma pack(pop)
static_assert(sizeof(DGraphHeader) == 20, "packed .dgraph header");

int kz_dgraph_load(const char* path, kz_csr* out) {
    if (!path || !out) KZ_FAIL(KZ_ERR_INVALID, "null argument");
    int fd = open(path, O_RDONLY);
    if (fd < 0) KZ_FAIL(KZ_ERR_IO, "cannot open %s", path);
    struct stat st;
    if (fstat(fd, &st) != 0 || size_t(st.st_size) < sizeof(DGraphHeader)) {
        close(fd);
        KZ_FAIL(KZ_ERR_IO, "%s: too short for a .dgraph header", path);
    }
    size_t len = size_t(st.st_size);
    void* map = mmap(nullptr, len, PROT_READ, MAP_PRIVATE, fd, 0);
    close(fd);
    if (map == MAP_FAILED) KZ_FAIL(KZ_ERR_IO, "mmap %s failed", path);
    DGraphHeader h;
    std::memcpy(&h, map, sizeof(h));
    auto fail = [&](int code, const char* msg) {
        munmap(map, len);
        kz_csr_release(out);
        kz::set_error("%s: %s", path, msg);
        return code;
    };
    if (h.has_edge_label) return fail(KZ_ERR_UNSUPPORTED, "edge labels are not supported");
    size_t lab_bytes = h.has_node_label ? sizeof(uint32_t) * size_t(h.num_vertices) : 0;
    size_t need = sizeof(DGraphHeader) + lab_bytes + size_t(h.num_edges) * 8;
    if (need != len) return fail(KZ_ERR_IO, "file size does not match header");
    if (!csr_alloc(out, h.num_vertices, h.num_edges, h.has_node_label != 0))
        return fail(KZ_ERR_NOMEM, "csr allocation");
    const uint8_t* base = static_cast<const uint8_t*>(map);
    if (lab_bytes) std::memcpy(out->labels, base + sizeof(DGraphHeader), lab_bytes);
    const uint32_t* ed = reinterpret_cast<const uint32_t*>(base + sizeof(DGraphHeader) + lab_bytes);
    const uint32_t n = h.num_vertices;
    const uint64_t m = h.num_edges;
    // Edges are sorted by (src, dst): row offsets come from the first edge of
    // each source (graph_loader.h:131-141); empty rows take the next start.
    int bad = 0, oriented = 1;
#pragma omp parallel for schedule(static) reduction(| : bad) reduction(& : oriented)
    for (uint64_t e = 0; e < m; ++e) {
        uint32_t s = ed[2 * e], d = ed[2 * e + 1];
        if (s >= n || d >= n) {
            bad |= 1;
            continue;
        }
        if (e > 0) {
            uint32_t ps = ed[2 * e - 2], pd = ed[2 * e - 1];
            if (ps > s || (ps == s && pd >= d)) bad |= 2;
        }
        if (d <= s) oriented = 0;
        out->col[e] = d;
    }
    if (bad & 1) return fail(KZ_ERR_IO, "vertex id out of range");
    if (bad & 2) return fail(KZ_ERR_IO, "edges not strictly sorted by (src, dst)");
    std::vector<uint32_t> deg(n, 0);
#pragma omp parallel for schedule(static)
    for (uint64_t e = 0; e < m; ++e) {
        // sorted input: count run lengths by detecting run starts
        uint32_t s = ed[2 * e];
        if (e == 0 || ed[2 * e - 2] != s) {
            uint64_t f = e;
            while (f < m && ed[2 * f] == s) ++f;
            deg[s] = uint32_t(f - e);
        }
    }
    exclusive_scan(deg.data(), out->rowptr, n);
    uint32_t maxd = 0;
    for (uint32_t u = 0; u < n; ++u) maxd = std::max(maxd, deg[u]);
    out->max_degree = maxd;
    out->oriented = (m > 0 && oriented) ? 1u : 0u;
    munmap(map, len);
    if (out->rowptr[n] != m) {
        kz_csr_release(out);
        KZ_FAIL(KZ_ERR_IO, "%s: degree sum mismatch", path);
    }
    return KZ_OK;
}

int kz_dgraph_store(const char* path, const kz_csr* g) {
    if (!path || !g || !g->rowptr) KZ_FAIL(KZ_ERR_INVALID, "null argument");
    FILE* f = std::fopen(path, "wb");
    if (!f) KZ_FAIL(KZ_ERR_IO, "cannot create %s", path);
    DGraphHeader h;
    h.num_vertices = g->num_vertices;
    h.num_edges = g->num_entries;
    h.has_node_label = g->has_labels && g->labels ? 1 : 0;
    h.has_edge_label = 0;
    bool ok = std::fwrite(&h, sizeof(h), 1, f) == 1;
    if (ok && h.has_node_label)
        ok = std::fwrite(g->labels, sizeof(uint32_t), g->num_vertices, f) == g->num_vertices;
    const size_t chunk = size_t(1) << 22;
    std::vector<uint32_t> buf(2 * chunk);
    uint32_t u = 0;
    for (uint64_t e0 = 0; ok && e0 < g->num_entries; e0 += chunk) {
        uint64_t e1 = std::min<uint64_t>(g->num_entries, e0 + chunk);
        for (uint64_t e = e0; e < e1; ++e) {
            while (g->rowptr[u + 1] <= e) ++u;
            buf[2 * (e - e0)] = u;
            buf[2 * (e - e0) + 1] = g->col[e];
        }
        ok = std::fwrite(buf.data(), 8, size_t(e1 - e0), f) == size_t(e1 - e0);
    }
    if (std::fclose(f) != 0) ok = false;
    if (!ok) KZ_FAIL(KZ_ERR_IO, "write to %s failed", path);
    return KZ_OK;
}

}  // extern "C"
