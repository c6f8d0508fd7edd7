// kz_oracle — CPU restatement of the OpenKhuzdulGraph embedding-extension
// path, used ONLY as test infrastructure: the checker in tests/, the
// cpu_baseline leg of bench.py and the smoke check.  Nothing in the product
// (openkhuzdulgraph_amd/) links, loads or calls this library.
//
// What it restates (paths relative to the OpenKhuzdulGraph source tree):
//   * set primitives: src/vertex_set.cc:20-184 (intersect_with,
//     subtracted_by, intersect_and_count, bounded_*), include/utilities.h:
//     64-86 (trim_v_list, trim_v_list_lower_bound);
//   * the engine's extension semantics: every vertex is a root
//     (src/distributed_application.cc:314-328), an embedding of size s is
//     passed to the app's apply(), scatter() makes a size s+1 child
//     (src/engine.cc:1385-1453); the level-by-level BFS order of the
//     reference does not change counts, so the restatement walks the same
//     tree depth-first;
//   * each app's generated apply() schedule, cited per function below;
//   * the labelled FSM app kautomine/three_fsm.cc (mining_frequent_patterns,
//     :688-695): its candidate generators (:276-604), apply() (:52-250) and
//     DistributedPatternDomain's MNI support (include/engine.h:405-580);
//   * Aggregator<EdgeId>: a uint64 sum wrapping mod 2^64
//     (include/engine.h:615-673), divided by the app's multiplicity
//     (kautomine/three_motif.cc:44, :280-284).
// Pinning (how this restatement is itself checked) lives in
// tests/test_oracle.py: std::set_* / brute force / closed forms / the
// reference binaries built by oracle/Makefile into oracle/_ref.

#include <omp.h>

#include "kz.h"  // plan IR layout only

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <vector>

typedef uint32_t VertexId;
typedef uint64_t EdgeId;

extern "C" {

// ---------------------------------------------------------------- a1..a3 --

// vertex_set.cc:20-49.  Early-out on an empty operand, then a merge that
// advances the operand cursor past smaller ids and emits equal ids.
uint32_t kzo_intersect(const VertexId* a, uint32_t na, const VertexId* b, uint32_t nb, VertexId* out) {
    if (na == 0 || nb == 0) return 0;
    uint32_t j = 0, k = 0;
    for (uint32_t i = 0; i < na; ++i) {
        VertexId v = a[i];
        while (j < nb && b[j] < v) ++j;
        if (j < nb && b[j] == v) {
            out[k++] = v;
            ++j;
        }
    }
    return k;
}

// vertex_set.cc:51-86: elements of a not found in b.
uint32_t kzo_subtract(const VertexId* a, uint32_t na, const VertexId* b, uint32_t nb, VertexId* out) {
    if (na == 0) return 0;
    uint32_t j = 0, k = 0;
    for (uint32_t i = 0; i < na; ++i) {
        VertexId v = a[i];
        bool found = false;
        while (j < nb) {
            if (b[j] == v) {
                found = true;
                ++j;
                break;
            }
            if (b[j] > v) break;
            ++j;
        }
        if (!found) out[k++] = v;
    }
    return k;
}

// vertex_set.cc:88-111.
uint32_t kzo_intersect_count(const VertexId* a, uint32_t na, const VertexId* b, uint32_t nb) {
    if (na == 0 || nb == 0) return 0;
    uint32_t j = 0, k = 0;
    for (uint32_t i = 0; i < na; ++i) {
        VertexId v = a[i];
        while (j < nb && b[j] < v) ++j;
        if (j < nb && b[j] == v) {
            ++k;
            ++j;
        }
    }
    return k;
}

// vertex_set.cc:113-145: stops at the first element of a >= upper_bound.
uint32_t kzo_bounded_intersect(const VertexId* a, uint32_t na, const VertexId* b, uint32_t nb,
                               VertexId upper_bound, VertexId* out) {
    if (na == 0 || nb == 0) return 0;
    uint32_t j = 0, k = 0;
    for (uint32_t i = 0; i < na; ++i) {
        VertexId v = a[i];
        if (v >= upper_bound) break;
        while (j < nb && b[j] < v) ++j;
        if (j < nb && b[j] == v) {
            out[k++] = v;
            ++j;
        }
    }
    return k;
}

// vertex_set.cc:147-184.
uint32_t kzo_bounded_subtract(const VertexId* a, uint32_t na, const VertexId* b, uint32_t nb,
                              VertexId upper_bound, VertexId* out) {
    if (na == 0) return 0;
    uint32_t j = 0, k = 0;
    for (uint32_t i = 0; i < na; ++i) {
        VertexId v = a[i];
        if (v >= upper_bound) break;
        bool found = false;
        while (j < nb) {
            if (b[j] == v) {
                found = true;
                ++j;
                break;
            }
            if (b[j] > v) break;
            ++j;
        }
        if (!found) out[k++] = v;
    }
    return k;
}

// utilities.h:64-74: number of leading elements < upper_bound.
uint32_t kzo_trim(const VertexId* v, uint32_t n, VertexId upper_bound) {
    if (n == 0 || v[0] >= upper_bound) return 0;
    if (v[n - 1] < upper_bound) return n;
    uint32_t l = 0, r = n;
    while (r - l > 1) {
        uint32_t m = (l + r) >> 1;
        if (v[m] < upper_bound) l = m;
        else r = m;
    }
    return l + 1;
}

// utilities.h:76-86: index of the first element > lower_bound (n if none).
uint32_t kzo_trim_lower_bound(const VertexId* v, uint32_t n, VertexId lower_bound) {
    if (n == 0 || v[n - 1] <= lower_bound) return n;
    if (v[0] > lower_bound) return 0;
    uint32_t l = 0, r = n;
    while (r - l > 1) {
        uint32_t m = (l + r) >> 1;
        if (v[m] <= lower_bound) l = m;
        else r = m;
    }
    return l + 1;
}

}  // extern "C"

// ------------------------------------------------------------ app schedules --

namespace {

struct G {
    const uint64_t* rp;
    const VertexId* col;
    uint32_t n;
    const VertexId* nbr(VertexId v) const { return col + rp[v]; }
    uint32_t deg(VertexId v) const { return uint32_t(rp[v + 1] - rp[v]); }
};

// kgraphpi's last-level count: |{v in L : v > lb}| minus every matched
// vertex found above lb (kgraphpi/triangle.cc:100-107 and siblings).
inline EdgeId kgraphpi_tail(const VertexId* L, uint32_t n, VertexId lb, const VertexId* m, int nm) {
    uint32_t t = kzo_trim_lower_bound(L, n, lb);
    EdgeId c = EdgeId(n - t);
    for (int i = 0; i < nm; ++i) c -= std::binary_search(L + t, L + n, m[i]);
    return c;
}

struct Buf {
    std::vector<VertexId> p[5];
    void ensure(uint32_t maxd) {
        for (auto& x : p)
            if (x.size() < maxd + 1) x.resize(maxd + 1);
    }
};

// kautomine/triangle.cc:58-121.
EdgeId ka_triangle(const G& g, VertexId v0, Buf& b) {
    EdgeId c = 0;
    const VertexId* n0 = g.nbr(v0);
    uint32_t d0 = g.deg(v0);
    for (uint32_t i = 0; i < d0; ++i) {
        VertexId v1 = n0[i];
        if (!(v1 > v0)) continue;                                      // :79
        uint32_t k = kzo_intersect(n0, d0, g.nbr(v1), g.deg(v1), b.p[1].data());  // :98
        for (uint32_t j = 0; j < k; ++j) c += b.p[1][j] > v1;          // :103-106
    }
    return c;
}

// kautomine/four_clique.cc:58-154 (prefix_1 cached at size 2, last level
// intersects it with N(v_2) and counts v > v_2).
EdgeId ka_four_clique(const G& g, VertexId v0, Buf& b) {
    EdgeId c = 0;
    const VertexId* n0 = g.nbr(v0);
    uint32_t d0 = g.deg(v0);
    for (uint32_t i = 0; i < d0; ++i) {
        VertexId v1 = n0[i];
        if (!(v1 > v0)) continue;                                            // :76
        uint32_t k1 = kzo_intersect(n0, d0, g.nbr(v1), g.deg(v1), b.p[1].data());  // :102
        for (uint32_t j = 0; j < k1; ++j) {
            VertexId v2 = b.p[1][j];
            if (!(v2 > v1)) continue;                                        // :108
            uint32_t k2 = kzo_intersect(b.p[1].data(), k1, g.nbr(v2), g.deg(v2), b.p[2].data());  // :130
            for (uint32_t t = 0; t < k2; ++t) c += b.p[2][t] > v2;           // :134-138
        }
    }
    return c;
}

// kautomine/five_clique.cc:58-187.
EdgeId ka_five_clique(const G& g, VertexId v0, Buf& b) {
    EdgeId c = 0;
    const VertexId* n0 = g.nbr(v0);
    uint32_t d0 = g.deg(v0);
    for (uint32_t i = 0; i < d0; ++i) {
        VertexId v1 = n0[i];
        if (!(v1 > v0)) continue;
        uint32_t k1 = kzo_intersect(n0, d0, g.nbr(v1), g.deg(v1), b.p[1].data());  // :102
        for (uint32_t j = 0; j < k1; ++j) {
            VertexId v2 = b.p[1][j];
            if (!(v2 > v1)) continue;                                        // :108
            uint32_t k2 = kzo_intersect(b.p[1].data(), k1, g.nbr(v2), g.deg(v2), b.p[2].data());  // :134
            for (uint32_t t = 0; t < k2; ++t) {
                VertexId v3 = b.p[2][t];
                if (!(v3 > v2)) continue;                                    // :140
                uint32_t k3 = kzo_intersect(b.p[2].data(), k2, g.nbr(v3), g.deg(v3), b.p[3].data());  // :163
                for (uint32_t q = 0; q < k3; ++q) c += b.p[3][q] > v3;      // :167-171
            }
        }
    }
    return c;
}

// kautomine/three_motif.cc:59-139, fused patterns (ENABLE_COMPUTE_FUSION):
// counts[0] = triangles (v_0 < v_1 guard, v > v_1), counts[1] = wedges
// from N(v_0) \ N(v_1) with v != v_1 (multiplicity 2).
void ka_three_motif(const G& g, VertexId v0, Buf& b, EdgeId* c) {
    const VertexId* n0 = g.nbr(v0);
    uint32_t d0 = g.deg(v0);
    for (uint32_t i = 0; i < d0; ++i) {
        VertexId v1 = n0[i];                                                 // :79 valid = true
        if (v0 < v1) {                                                       // :102
            uint32_t k = kzo_intersect(n0, d0, g.nbr(v1), g.deg(v1), b.p[1].data());  // :104
            for (uint32_t j = 0; j < k; ++j) c[0] += b.p[1][j] > v1;        // :106-110
        }
        uint32_t k = kzo_subtract(n0, d0, g.nbr(v1), g.deg(v1), b.p[2].data());  // :115
        for (uint32_t j = 0; j < k; ++j) c[1] += b.p[2][j] != v1;           // :117-121
    }
}

// kgraphpi/triangle.cc:53-122.
EdgeId kg_triangle(const G& g, VertexId v0, Buf& b) {
    EdgeId c = 0;
    const VertexId* n0 = g.nbr(v0);
    uint32_t d0 = g.deg(v0);
    for (uint32_t i = 0; i < d0; ++i) {
        VertexId v1 = n0[i];
        if (v1 <= v0) continue;                                             // :72
        if (!(v1 != v0)) continue;                                          // :74
        uint32_t k = kzo_intersect(n0, d0, g.nbr(v1), g.deg(v1), b.p[1].data());  // :98
        VertexId m[2] = {v0, v1};
        c += kgraphpi_tail(b.p[1].data(), k, v1, m, 2);                     // :100-107
    }
    return c;
}

// kgraphpi/four_clique.cc:53-159.
EdgeId kg_four_clique(const G& g, VertexId v0, Buf& b) {
    EdgeId c = 0;
    const VertexId* n0 = g.nbr(v0);
    uint32_t d0 = g.deg(v0);
    for (uint32_t i = 0; i < d0; ++i) {
        VertexId v1 = n0[i];
        if (v1 <= v0 || !(v1 != v0)) continue;                              // :73-76
        uint32_t k1 = kzo_intersect(n0, d0, g.nbr(v1), g.deg(v1), b.p[1].data());  // :99
        for (uint32_t j = 0; j < k1; ++j) {
            VertexId v2 = b.p[1][j];
            if (v2 <= v1) continue;                                         // :105
            if (!(v2 != v0 && v2 != v1)) continue;                          // :107
            uint32_t k2 = kzo_intersect(b.p[1].data(), k1, g.nbr(v2), g.deg(v2), b.p[2].data());  // :133
            VertexId m[3] = {v0, v1, v2};
            c += kgraphpi_tail(b.p[2].data(), k2, std::max(v0, v2), m, 3);  // :135-144
        }
    }
    return c;
}

// kgraphpi/five_clique.cc:53-195.
EdgeId kg_five_clique(const G& g, VertexId v0, Buf& b) {
    EdgeId c = 0;
    const VertexId* n0 = g.nbr(v0);
    uint32_t d0 = g.deg(v0);
    for (uint32_t i = 0; i < d0; ++i) {
        VertexId v1 = n0[i];
        if (v1 <= v0 || !(v1 != v0)) continue;                              // :68-71
        uint32_t k1 = kzo_intersect(n0, d0, g.nbr(v1), g.deg(v1), b.p[1].data());  // :99
        for (uint32_t j = 0; j < k1; ++j) {
            VertexId v2 = b.p[1][j];
            if (v2 <= v1 || !(v2 != v0 && v2 != v1)) continue;              // :103-107
            uint32_t k2 = kzo_intersect(b.p[1].data(), k1, g.nbr(v2), g.deg(v2), b.p[2].data());  // :133
            for (uint32_t t = 0; t < k2; ++t) {
                VertexId v3 = b.p[2][t];
                if (v3 <= v0 || v3 <= v2) continue;                         // :139
                if (!(v3 != v0 && v3 != v1 && v3 != v2)) continue;          // :141
                uint32_t k3 = kzo_intersect(b.p[2].data(), k2, g.nbr(v3), g.deg(v3), b.p[3].data());  // :168
                VertexId m[4] = {v0, v1, v2, v3};
                c += kgraphpi_tail(b.p[3].data(), k3, std::max(v1, v3), m, 4);  // :170-180
            }
        }
    }
    return c;
}

// kgraphpi/three_motif.cc:52-174, two patterns run one after the other.
// Pattern 0 (:56-108): triangles with the kgraphpi tail count.
// Pattern 1 (:111-166): wedges centred at v_0; v_1 is scattered inactive
// and finalized (:140), the tail count runs over N(v_0) with lb = v_1.
EdgeId kg_three_motif_p1(const G& g, VertexId v0) {
    EdgeId c = 0;
    const VertexId* n0 = g.nbr(v0);
    uint32_t d0 = g.deg(v0);
    for (uint32_t i = 0; i < d0; ++i) {
        VertexId v1 = n0[i];
        if (!(v1 != v0)) continue;                                          // :135-137
        VertexId m[2] = {v0, v1};
        c += kgraphpi_tail(n0, d0, v1, m, 2);                               // :153-161
    }
    return c;
}

}  // namespace

extern "C" {

// App identifiers accepted by kzo_run_app.
enum {
    KZO_KA_TRIANGLE = 0,
    KZO_KA_FOUR_CLIQUE = 1,
    KZO_KA_FIVE_CLIQUE = 2,
    KZO_KA_THREE_MOTIF = 3,
    KZO_KG_TRIANGLE = 4,
    KZO_KG_FOUR_CLIQUE = 5,
    KZO_KG_FIVE_CLIQUE = 6,
    KZO_KG_THREE_MOTIF = 7,
};

// Runs one app over the roots `roots[0..nroots)` (all vertices when roots is
// NULL and nroots == n).  Writes the raw aggregator sums (before the
// multiplicity division) to raw[] and the printed counts to out[]; returns
// the number of patterns, or -1 for an unknown app.
int kzo_run_app(int app, const uint64_t* rowptr, const uint32_t* col, uint32_t n,
                const uint32_t* roots, uint64_t nroots, int threads, uint64_t* raw, uint64_t* out) {
    G g{rowptr, col, n};
    int np = (app == KZO_KA_THREE_MOTIF || app == KZO_KG_THREE_MOTIF) ? 2 : 1;
    EdgeId acc[2] = {0, 0};
    auto run1 = [&](EdgeId (*fn)(const G&, VertexId, Buf&)) {
        EdgeId total = 0;
        uint32_t maxd = 0;
        for (uint32_t v = 0; v < g.n; ++v) maxd = std::max(maxd, g.deg(v));
        int T = threads > 0 ? threads : omp_get_max_threads();
#pragma omp parallel num_threads(T) reduction(+ : total)
        {
            Buf b;
            b.ensure(maxd);
#pragma omp for schedule(dynamic, 64)
            for (uint64_t i = 0; i < nroots; ++i) total += fn(g, roots ? roots[i] : VertexId(i), b);
        }
        return total;
    };
    switch (app) {
        case KZO_KA_TRIANGLE: acc[0] = run1(ka_triangle); break;
        case KZO_KA_FOUR_CLIQUE: acc[0] = run1(ka_four_clique); break;
        case KZO_KA_FIVE_CLIQUE: acc[0] = run1(ka_five_clique); break;
        case KZO_KG_TRIANGLE: acc[0] = run1(kg_triangle); break;
        case KZO_KG_FOUR_CLIQUE: acc[0] = run1(kg_four_clique); break;
        case KZO_KG_FIVE_CLIQUE: acc[0] = run1(kg_five_clique); break;
        case KZO_KA_THREE_MOTIF: {
            EdgeId c0 = 0, c1 = 0;
            uint32_t maxd = 0;
            for (uint32_t v = 0; v < g.n; ++v) maxd = std::max(maxd, g.deg(v));
            int T = threads > 0 ? threads : omp_get_max_threads();
#pragma omp parallel num_threads(T) reduction(+ : c0, c1)
            {
                Buf b;
                b.ensure(maxd);
#pragma omp for schedule(dynamic, 64)
                for (uint64_t i = 0; i < nroots; ++i) {
                    EdgeId c[2] = {0, 0};
                    ka_three_motif(g, roots ? roots[i] : VertexId(i), b, c);
                    c0 += c[0];
                    c1 += c[1];
                }
            }
            acc[0] = c0;
            acc[1] = c1;
            break;
        }
        case KZO_KG_THREE_MOTIF: {
            acc[0] = run1(kg_triangle);
            EdgeId c1 = 0;
            int T = threads > 0 ? threads : omp_get_max_threads();
#pragma omp parallel for num_threads(T) schedule(dynamic, 64) reduction(+ : c1)
            for (uint64_t i = 0; i < nroots; ++i) c1 += kg_three_motif_p1(g, roots ? roots[i] : VertexId(i));
            acc[1] = c1;
            break;
        }
        default: return -1;
    }
    // Multiplicities: kautomine three_motif {1, 2} (three_motif.cc:44);
    // every other app {1} (kgraphpi three_motif has none, :231-256).
    for (int p = 0; p < np; ++p) {
        raw[p] = acc[p];
        EdgeId mult = (app == KZO_KA_THREE_MOTIF && p == 1) ? 2 : 1;
        out[p] = acc[p] / mult;
    }
    return np;
}

// ------------------------------------------------------------ brute force --

// ------------------------------------------------------- plan interpreter --
// Depth-first CPU interpreter of the plan IR (include/kz.h), restating the
// engine semantics a plan encodes: every vertex is a root; an embedding of
// size s runs the step list of its order code; set operations follow
// vertex_set.cc (kzo_intersect / kzo_subtract above); COUNT adds the number
// of candidates whose position-class bit is set, SCATTER recurses on them.

namespace {

struct PlanRun {
    const kz_plan_desc* p;
    G g;
    uint64_t acc[KZ_MAX_AGGREGATORS];
    // slot storage per size and op index for the current DFS path
    std::vector<VertexId> slot[KZ_MAX_PATTERN + 1][KZ_MAX_OPS];
    uint32_t slot_len[KZ_MAX_PATTERN + 1][KZ_MAX_OPS];
    uint32_t maxd;

    void init(uint32_t md) {
        maxd = md;
        for (auto& row : slot)
            for (auto& v : row) v.assign(md + 1, 0);
        std::memset(acc, 0, sizeof(acc));
    }

    static uint32_t code_of(const VertexId* m, int s) {
        uint32_t code = 0, pw = 1;
        for (int j = 0; j < s; ++j) {
            uint32_t r = 0;
            for (int i = 0; i < s; ++i) {
                if (!(m[i] < m[j])) continue;
                bool first = true;
                for (int k = 0; k < i; ++k)
                    if (m[k] == m[i]) first = false;
                r += first;
            }
            code += r * pw;
            pw *= uint32_t(s);
        }
        return code;
    }

    void set_of(const kz_set_ref& r, const VertexId* m, const VertexId** ptr, uint32_t* n) {
        if (r.kind == KZ_SET_NBR) {
            *ptr = g.nbr(m[r.a]);
            *n = g.deg(m[r.a]);
        } else {
            *ptr = slot[r.a][r.b].data();
            *n = slot_len[r.a][r.b];
        }
    }

    void extend(VertexId* m, int s) {
        if (s < 1 || s >= KZ_MAX_PATTERN || s >= int(p->pattern_size)) return;
        const uint8_t li = p->type_table[s][code_of(m, s)];
        if (li == 0xFF) return;
        const kz_step_list& L = p->lists[li];
        for (int j = 0; j < L.num_ops; ++j) {
            const VertexId *a, *b;
            uint32_t na, nb;
            set_of(L.ops[j].lhs, m, &a, &na);
            set_of(L.ops[j].rhs, m, &b, &nb);
            slot_len[s][j] = L.ops[j].op == KZ_OP_INTERSECT ? kzo_intersect(a, na, b, nb, slot[s][j].data())
                                                            : kzo_subtract(a, na, b, nb, slot[s][j].data());
        }
        VertexId z[KZ_MAX_PATTERN];
        int q = 0;
        for (int i = 0; i < s; ++i) z[q++] = m[i];
        std::sort(z, z + q);
        q = int(std::unique(z, z + q) - z);
        for (int o = 0; o < L.num_outs; ++o) {
            const kz_out& out = L.outs[o];
            const VertexId* src;
            uint32_t n;
            set_of(out.src, m, &src, &n);
            // copy: a child may overwrite slots of this size further down
            std::vector<VertexId> cand(src, src + n);
            for (VertexId v : cand) {
                const uint32_t cls = 2 * uint32_t(std::lower_bound(z, z + q, v) - z) +
                                     (std::binary_search(z, z + q, v) ? 1u : 0u);
                if (!((out.pred >> cls) & 1u)) continue;
                if (out.kind == KZ_OUT_COUNT) {
                    acc[out.aggregator] += 1;
                } else {
                    m[s] = v;
                    extend(m, s + 1);
                }
            }
        }
    }
};

}  // namespace

// raw[a] = uint64 sum of aggregator a over all roots.
int kzo_run_plan(const kz_plan_desc* plan, const uint64_t* rowptr, const uint32_t* col, uint32_t n, int threads,
                 uint64_t* raw) {
    uint32_t maxd = 0;
    for (uint32_t v = 0; v < n; ++v) maxd = std::max(maxd, uint32_t(rowptr[v + 1] - rowptr[v]));
    uint64_t tot[KZ_MAX_AGGREGATORS] = {0};
    int T = threads > 0 ? threads : omp_get_max_threads();
#pragma omp parallel num_threads(T)
    {
        PlanRun r;
        r.p = plan;
        r.g = G{rowptr, col, n};
        r.init(maxd);
#pragma omp for schedule(dynamic, 64)
        for (uint64_t v = 0; v < n; ++v) {
            VertexId m[KZ_MAX_PATTERN + 1];
            m[0] = VertexId(v);
            r.extend(m, 1);
        }
#pragma omp critical
        for (int a = 0; a < KZ_MAX_AGGREGATORS; ++a) tot[a] += r.acc[a];
    }
    for (uint32_t a = 0; a < plan->num_aggregators && a < KZ_MAX_AGGREGATORS; ++a) raw[a] = tot[a];
    return 0;
}

// Exhaustive k-subset enumeration (k = 3..5) with an adjacency matrix, for
// n <= 256.  Counts k-cliques; pins the app restatements by closed form
// independent of any intersection code.
long long kzo_brute_cliques(const uint64_t* rowptr, const uint32_t* col, uint32_t n, int k) {
    if (n > 256 || k < 3 || k > 5) return -1;
    std::vector<uint8_t> A(size_t(n) * n, 0);
    for (uint32_t u = 0; u < n; ++u)
        for (uint64_t e = rowptr[u]; e < rowptr[u + 1]; ++e) {
            A[size_t(u) * n + col[e]] = 1;
            A[size_t(col[e]) * n + u] = 1;
        }
    auto a = [&](uint32_t x, uint32_t y) { return A[size_t(x) * n + y] != 0; };
    long long c = 0;
    for (uint32_t x0 = 0; x0 < n; ++x0)
        for (uint32_t x1 = x0 + 1; x1 < n; ++x1) {
            if (!a(x0, x1)) continue;
            for (uint32_t x2 = x1 + 1; x2 < n; ++x2) {
                if (!a(x0, x2) || !a(x1, x2)) continue;
                if (k == 3) {
                    ++c;
                    continue;
                }
                for (uint32_t x3 = x2 + 1; x3 < n; ++x3) {
                    if (!a(x0, x3) || !a(x1, x3) || !a(x2, x3)) continue;
                    if (k == 4) {
                        ++c;
                        continue;
                    }
                    for (uint32_t x4 = x3 + 1; x4 < n; ++x4)
                        c += a(x0, x4) && a(x1, x4) && a(x2, x4) && a(x3, x4);
                }
            }
        }
    return c;
}

// Induced connected 3- and 4-vertex subgraph census by exhaustive subset
// enumeration on a symmetric graph with n <= 128.  out[0..2) = {wedge
// (open 2-path), triangle}; out[2..8) = {3-path, 3-star, 4-cycle, tailed
// triangle, diamond, 4-clique}.  Classification is by the sorted degree
// sequence of the induced subgraph, which separates all six connected
// 4-vertex graphs.
int kzo_brute_motifs(const uint64_t* rowptr, const uint32_t* col, uint32_t n, uint64_t* out) {
    if (n > 128) return -1;
    std::vector<uint8_t> A(size_t(n) * n, 0);
    for (uint32_t u = 0; u < n; ++u)
        for (uint64_t e = rowptr[u]; e < rowptr[u + 1]; ++e) {
            A[size_t(u) * n + col[e]] = 1;
            A[size_t(col[e]) * n + u] = 1;
        }
    auto a = [&](uint32_t x, uint32_t y) { return int(A[size_t(x) * n + y]); };
    for (int i = 0; i < 8; ++i) out[i] = 0;
    for (uint32_t x = 0; x < n; ++x)
        for (uint32_t y = x + 1; y < n; ++y)
            for (uint32_t z = y + 1; z < n; ++z) {
                int e = a(x, y) + a(x, z) + a(y, z);
                if (e == 2) out[0]++;
                if (e == 3) out[1]++;
                for (uint32_t w = z + 1; w < n; ++w) {
                    int d[4] = {a(x, y) + a(x, z) + a(x, w), a(x, y) + a(y, z) + a(y, w),
                                a(x, z) + a(y, z) + a(z, w), a(x, w) + a(y, w) + a(z, w)};
                    int m = (d[0] + d[1] + d[2] + d[3]) / 2;
                    std::sort(d, d + 4);
                    if (d[0] == 0) continue;  // a vertex with no edge: disconnected
                    if (m == 3) {
                        if (d[3] == 3) out[3]++;                   // star 1,1,1,3
                        else if (d[0] == 1 && d[1] == 1) out[2]++;  // path 1,1,2,2
                        // 0,2,2,2 (triangle + isolated) excluded by d[0]==0
                    } else if (m == 4) {
                        if (d[0] == 2) out[4]++;  // cycle 2,2,2,2
                        else out[5]++;            // paw 1,2,2,3
                    } else if (m == 5) {
                        out[6]++;
                    } else if (m == 6) {
                        out[7]++;
                    }
                }
            }
    return 8;
}

// Merge-model algorithmic bytes of the k-clique schedule on an oriented
// graph: every intersection reads both operands once, every materialised
// prefix is written once, 4 bytes per id.
//   B3 = 4 * [ m + sum_{(u,v)} (d+u + d+v) ]
//   B4 = B3 + 4 * [ T + sum_{(u,v)} c(u,v)^2 + sum_{(u,v,w) in Tri} d+w ]
//   B5 = B4 + 4 * [ K4 + sum_{(u,v,w)} c(u,v,w)^2 + sum_{K4 (u,v,w,x)} d+x ]
// out[0..3) = {B3, B4, B5}; out[3..6) = {triangles, 4-cliques, 5-cliques}.
// With roots != NULL only the work of those root vertices is counted (every
// term of B belongs to the root u of its edge (u, v)).
int kzo_clique_bytes(const uint64_t* rowptr, const uint32_t* col, uint32_t n, int kmax, int threads,
                     const uint32_t* roots, uint64_t nroots, uint64_t* out) {
    G g{rowptr, col, n};
    uint64_t m = 0;
    uint64_t s3 = 0, T = 0, s4 = 0, K4 = 0, s5 = 0, K5 = 0;
    uint32_t maxd = 0;
    for (uint32_t v = 0; v < n; ++v) maxd = std::max(maxd, g.deg(v));
    if (!roots) nroots = n;
    int nt = threads > 0 ? threads : omp_get_max_threads();
#pragma omp parallel num_threads(nt) reduction(+ : m, s3, T, s4, K4, s5, K5)
    {
        std::vector<VertexId> p1(maxd + 1), p2(maxd + 1);
#pragma omp for schedule(dynamic, 64)
        for (uint64_t r = 0; r < nroots; ++r) {
            const uint64_t u = roots ? roots[r] : r;
            m += g.deg(VertexId(u));
            for (uint64_t e = rowptr[u]; e < rowptr[u + 1]; ++e) {
                VertexId v = col[e];
                s3 += g.deg(VertexId(u)) + g.deg(v);
                if (kmax < 4) continue;
                uint32_t c = kzo_intersect(g.nbr(VertexId(u)), g.deg(VertexId(u)), g.nbr(v), g.deg(v), p1.data());
                T += c;
                s4 += uint64_t(c) * c;
                for (uint32_t i = 0; i < c; ++i) {
                    VertexId w = p1[i];
                    s4 += g.deg(w);
                    if (kmax < 5) {
                        K4 += kzo_intersect_count(p1.data(), c, g.nbr(w), g.deg(w));
                        continue;
                    }
                    uint32_t c2 = kzo_intersect(p1.data(), c, g.nbr(w), g.deg(w), p2.data());
                    K4 += c2;
                    s5 += uint64_t(c2) * c2;
                    for (uint32_t j = 0; j < c2; ++j) {
                        VertexId x = p2[j];
                        s5 += g.deg(x);
                        K5 += kzo_intersect_count(p2.data(), c2, g.nbr(x), g.deg(x));
                    }
                }
            }
        }
    }
    out[0] = 4 * (m + s3);
    out[1] = out[0] + 4 * (T + s4);
    out[2] = out[1] + 4 * (K4 + s5);
    out[3] = T;
    out[4] = K4;
    out[5] = K5;
    return 0;
}

// ------------------------------------------------------------ motif census --
// Induced connected 3- and 4-vertex census of a symmetric graph (sorted
// adjacency, no loops or duplicates), in the order of kzo_brute_motifs:
// {wedge, triangle, 3-path, 3-star, 4-cycle, tailed triangle, diamond,
// 4-clique}.  The reference has no 4-vertex census app; this follows the
// standard combinatorial route.  Non-induced counts come from degrees d,
// per-edge triangle counts t(e) = |N(u) & N(v)|, per-vertex triangle counts
// t(v), 4-cycles and 4-cliques:
//   wedges = sum C(d,2)          3-stars = sum C(d,3)
//   3-paths = sum_e (d_u-1)(d_v-1) - 3T      tailed = sum_v t(v) (d_v - 2)
//   diamonds = sum_e C(t(e),2)   4-cycles: each counted once from its
//   highest-ranked vertex u (rank = (degree, id)) as C(c(w), 2) over the
//   vertices w reached by u -> v -> w with v, w ranked below u.
// Induced counts follow by back-substitution through the containment
// matrix of the six connected 4-vertex graphs (each row: copies of that
// graph inside 3-path, 3-star, 4-cycle, tailed, diamond, 4-clique):
//   3-path  1 0 4 2 6 12 | 3-star 0 1 0 1 2 4 | 4-cycle 0 0 1 0 1 3
//   tailed  0 0 0 1 4 12 | diamond 0 0 0 0 1 6 | 4-clique 0 0 0 0 0 1
// nonind (may be NULL) receives the non-induced counts in the same order.
// All sums wrap mod 2^64, so the differences are exact.
int kzo_motif_census(const uint64_t* rowptr, const uint32_t* col, uint32_t n, int threads, uint64_t* out,
                     uint64_t* nonind) {
    if (threads > 0) omp_set_num_threads(threads);
    auto deg = [&](uint32_t v) { return uint64_t(rowptr[v + 1] - rowptr[v]); };
    std::vector<uint32_t> order(n), rank(n);
    for (uint32_t v = 0; v < n; ++v) order[v] = v;
    std::sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) {
        return deg(a) != deg(b) ? deg(a) < deg(b) : a < b;
    });
    for (uint32_t i = 0; i < n; ++i) rank[order[i]] = i;
    // rank-oriented out-lists (kept in id order) for the 4-cliques
    std::vector<uint64_t> orp(size_t(n) + 1, 0);
    for (uint32_t v = 0; v < n; ++v) {
        uint64_t c = 0;
        for (uint64_t e = rowptr[v]; e < rowptr[v + 1]; ++e) c += rank[col[e]] > rank[v];
        orp[v + 1] = orp[v] + c;
    }
    std::vector<uint32_t> ocol(orp[n] ? orp[n] : 1);
#pragma omp parallel for schedule(dynamic, 1024)
    for (int64_t vi = 0; vi < int64_t(n); ++vi) {
        const uint32_t v = uint32_t(vi);
        uint64_t o = orp[v];
        for (uint64_t e = rowptr[v]; e < rowptr[v + 1]; ++e)
            if (rank[col[e]] > rank[v]) ocol[o++] = col[e];
    }
    std::vector<uint64_t> tv(n, 0);  // sum of t(e) over the edges at v (= 2 t(v))
    uint64_t S2 = 0, S3 = 0, P3 = 0, TE = 0, DIA = 0, C4 = 0, K4 = 0;
#pragma omp parallel reduction(+ : S2, S3, P3, TE, DIA, C4, K4)
    {
        std::vector<uint32_t> cnt(n, 0), touched, p1;
        touched.reserve(1024);
#pragma omp for schedule(dynamic, 256)
        for (int64_t ui = 0; ui < int64_t(n); ++ui) {
            const uint32_t u = uint32_t(ui);
            const uint64_t du = deg(u);
            S2 += du * (du - 1) / 2;
            S3 += du * (du - 1) * (du - 2) / 6;
            const uint32_t* Nu = col + rowptr[u];
            for (uint64_t e = rowptr[u]; e < rowptr[u + 1]; ++e) {
                const uint32_t v = col[e];
                if (v <= u) continue;
                const uint64_t dv = deg(v);
                const uint64_t t = kzo_intersect_count(Nu, uint32_t(du), col + rowptr[v], uint32_t(dv));
                TE += t;
                DIA += t * (t - 1) / 2;
                P3 += (du - 1) * (dv - 1);
#pragma omp atomic
                tv[u] += t;
#pragma omp atomic
                tv[v] += t;
            }
            // 4-cycles with u ranked highest
            for (uint64_t e = rowptr[u]; e < rowptr[u + 1]; ++e) {
                const uint32_t v = col[e];
                if (rank[v] > rank[u]) continue;
                for (uint64_t f = rowptr[v]; f < rowptr[v + 1]; ++f) {
                    const uint32_t w = col[f];
                    if (rank[w] >= rank[u]) continue;
                    if (cnt[w]++ == 0) touched.push_back(w);
                }
            }
            for (uint32_t w : touched) {
                C4 += uint64_t(cnt[w]) * (cnt[w] - 1) / 2;
                cnt[w] = 0;
            }
            touched.clear();
            // 4-cliques u < v < w < x in rank order
            const uint32_t* Ou = ocol.data() + orp[u];
            const uint32_t ou = uint32_t(orp[u + 1] - orp[u]);
            p1.resize(ou);
            for (uint32_t i = 0; i < ou; ++i) {
                const uint32_t v = Ou[i];
                const uint32_t c = kzo_intersect(Ou, ou, ocol.data() + orp[v], uint32_t(orp[v + 1] - orp[v]), p1.data());
                for (uint32_t j = 0; j < c; ++j) {
                    const uint32_t w = p1[j];
                    K4 += kzo_intersect_count(p1.data(), c, ocol.data() + orp[w], uint32_t(orp[w + 1] - orp[w]));
                }
            }
        }
    }
    uint64_t PAW = 0;
    for (uint32_t v = 0; v < n; ++v) PAW += (tv[v] / 2) * (deg(v) >= 2 ? deg(v) - 2 : 0);
    const uint64_t T = TE / 3;
    const uint64_t N[8] = {S2, T, P3 - 3 * T, S3, C4, PAW, DIA, K4};
    if (nonind)
        for (int i = 0; i < 8; ++i) nonind[i] = N[i];
    const uint64_t iK4 = N[7];
    const uint64_t iDia = N[6] - 6 * iK4;
    const uint64_t iPaw = N[5] - 4 * iDia - 12 * iK4;
    const uint64_t iC4 = N[4] - iDia - 3 * iK4;
    const uint64_t iStar = N[3] - iPaw - 2 * iDia - 4 * iK4;
    const uint64_t iP3 = N[2] - 4 * iC4 - 2 * iPaw - 6 * iDia - 12 * iK4;
    out[0] = S2 - 3 * T;
    out[1] = T;
    out[2] = iP3;
    out[3] = iStar;
    out[4] = iC4;
    out[5] = iPaw;
    out[6] = iDia;
    out[7] = iK4;
    return 8;
}


// ------------------------------------------------------------ three_fsm --
// Frequent 3/4-vertex labelled patterns with MNI support, restating
// kautomine/three_fsm.cc.  Levels: 0 vertex, 1 edge, 2 wedge, 3 triangle,
// 4 four-chain, 5 three-star.  counts[l] / sums[l] are what the app prints
// for level l ("Number of frequent ... patterns: X, support sum ...: Y";
// level 0 sums the support of every label, :260-269).  Every evaluated
// candidate is appended to cand (if not NULL, up to cap entries) as
// {level, l0, l1, l2, l3, support, frequent}; *ncand gets the total.
// Each candidate is enumerated literally as apply() does (roots are the
// vertices of label labels[0], engine.cc:1459-1465), and a domain keeps the
// distinct vertices inserted with enable = true at each pattern position
// (engine.h:465-475); the support is the smallest domain (:476-560).
}  // extern "C"

#include <set>

namespace {
struct FsmPat {
    int size;
    uint32_t l[4];
    bool operator<(const FsmPat& o) const {
        if (size != o.size) return size < o.size;
        for (int i = 0; i < size; ++i)
            if (l[i] != o.l[i]) return l[i] < o.l[i];
        return false;
    }
};
struct FsmCand {
    FsmPat canon;
    std::vector<FsmPat> autos;
};
FsmPat fp(uint32_t a, uint32_t b) { return FsmPat{2, {a, b, 0, 0}}; }
FsmPat fp(uint32_t a, uint32_t b, uint32_t c) { return FsmPat{3, {a, b, c, 0}}; }
FsmPat fp(uint32_t a, uint32_t b, uint32_t c, uint32_t d) { return FsmPat{4, {a, b, c, d}}; }
}  // namespace

extern "C" {

int kzo_fsm(const uint64_t* rowptr, const uint32_t* col, uint32_t n, const uint32_t* lab, uint32_t threshold,
            uint32_t* counts, uint64_t* sums, uint32_t* cand, uint32_t cap, uint32_t* ncand) {
    uint32_t L = 0;
    for (uint32_t v = 0; v < n; ++v) L = std::max(L, lab[v] + 1);
    std::vector<uint32_t> lc(L, 0);
    for (uint32_t v = 0; v < n; ++v) ++lc[lab[v]];
    std::vector<char> fv(L);
    counts[0] = 0;
    sums[0] = 0;
    for (uint32_t l = 0; l < L; ++l) {  // setup(), :255-275
        fv[l] = lc[l] >= threshold;
        counts[0] += fv[l];
        sums[0] += lc[l];
    }
    std::set<FsmPat> fe, fw;  // frequent edge / wedge tables (with automorphisms)
    uint32_t nc = 0;
    auto nbr = [&](uint32_t v) { return std::make_pair(col + rowptr[v], col + rowptr[v + 1]); };
    std::vector<std::vector<char>> dom(4, std::vector<char>(n));
    std::vector<uint32_t> dsz(4);
    auto ins = [&](int i, uint32_t v, bool en) {  // DistributedPatternDomain::insert
        if (en && !dom[i][v]) {
            dom[i][v] = 1;
            ++dsz[i];
        }
    };
    // one candidate: enumerate as apply() does, return the MNI support
    auto support = [&](int level, const FsmPat& p) -> uint32_t {
        for (int i = 0; i < 4; ++i) std::fill(dom[i].begin(), dom[i].end(), 0), dsz[i] = 0;
        for (uint32_t v0 = 0; v0 < n; ++v0) {
            if (lab[v0] != p.l[0]) continue;
            auto [b0, e0] = nbr(v0);
            if (level == 1) {  // Edge, :55-68
                for (auto a = b0; a != e0; ++a) {
                    const bool en = lab[*a] == p.l[1];
                    ins(0, v0, en), ins(1, *a, en);
                }
                continue;
            }
            for (auto a = b0; a != e0; ++a) {
                const uint32_t v1 = *a;
                if (lab[v1] != p.l[1]) continue;
                if (level == 2) {  // Wedge, :70-103
                    for (auto c = b0; c != e0; ++c) {
                        const bool en = lab[*c] == p.l[2] && *c != v1;
                        ins(0, v0, en), ins(1, v1, en), ins(2, *c, en);
                    }
                } else if (level == 3) {  // Triangle, :105-147
                    auto [b1, e1] = nbr(v1);
                    std::vector<uint32_t> t;
                    std::set_intersection(b0, e0, b1, e1, std::back_inserter(t));
                    for (uint32_t v2 : t) {
                        const bool en = lab[v2] == p.l[2];
                        ins(0, v0, en), ins(1, v1, en), ins(2, v2, en);
                    }
                } else if (level == 4) {  // FourChain, :149-197
                    auto [b1, e1] = nbr(v1);
                    for (auto c = b0; c != e0; ++c) {
                        const uint32_t v2 = *c;
                        if (lab[v2] != p.l[2] || v1 == v2) continue;
                        for (auto d = b1; d != e1; ++d) {
                            const uint32_t v3 = *d;
                            const bool en = lab[v3] == p.l[3] && v3 != v0 && v3 != v2;
                            ins(0, v0, en), ins(1, v1, en), ins(2, v2, en), ins(3, v3, en);
                        }
                    }
                } else {  // ThreeStar, :199-247
                    auto [b1, e1] = nbr(v1);
                    for (auto c = b1; c != e1; ++c) {
                        const uint32_t v2 = *c;
                        if (lab[v2] != p.l[2] || v0 == v2) continue;
                        for (auto d = b1; d != e1; ++d) {
                            const uint32_t v3 = *d;
                            const bool en = lab[v3] == p.l[3] && v3 != v0 && v3 != v2;
                            ins(0, v0, en), ins(1, v1, en), ins(2, v2, en), ins(3, v3, en);
                        }
                    }
                }
            }
        }
        uint32_t s = n;  // get_support, :476-560
        for (int i = 0; i < p.size; ++i) s = std::min(s, dsz[i]);
        return s;
    };
    auto evaluate = [&](int level, const std::vector<FsmCand>& cs, std::set<FsmPat>* table) {
        counts[level] = 0;
        sums[level] = 0;
        for (const FsmCand& c : cs) {
            const uint32_t s = support(level, c.canon);
            const bool fr = s >= threshold;
            if (cand && nc < cap) {
                uint32_t* o = cand + size_t(nc) * 7;
                o[0] = uint32_t(level);
                for (int i = 0; i < 4; ++i) o[1 + i] = i < c.canon.size ? c.canon.l[i] : 0u;
                o[5] = s;
                o[6] = fr;
            }
            ++nc;
            if (fr) {
                if (table) {
                    table->insert(c.canon);
                    for (const FsmPat& a : c.autos) table->insert(a);
                }
                ++counts[level];
                sums[level] += s;
            }
        }
    };
    auto has = [](const std::set<FsmPat>& t, const FsmPat& p) { return t.count(p) != 0; };
    std::vector<FsmCand> cs;
    // edges, :282-304
    for (uint32_t a = 0; a < L; ++a)
        for (uint32_t b = 0; b < L; ++b)
            if (fv[a] && fv[b] && a <= b) cs.push_back({fp(a, b), {fp(b, a)}});
    evaluate(1, cs, &fe);
    // wedges, :306-339 (center first)
    cs.clear();
    for (uint32_t a = 0; a < L; ++a) {
        if (!fv[a]) continue;
        for (uint32_t b = 0; b < L; ++b) {
            if (!has(fe, fp(a, b))) continue;
            for (uint32_t c = 0; c < L; ++c)
                if (has(fe, fp(a, c)) && b <= c) cs.push_back({fp(a, b, c), {fp(a, c, b)}});
        }
    }
    evaluate(2, cs, &fw);
    // triangles, :341-399 (the table of frequent triangles is not read later)
    cs.clear();
    for (uint32_t a = 0; a < L; ++a) {
        if (!fv[a]) continue;
        for (uint32_t b = 0; b < L; ++b) {
            if (!has(fe, fp(a, b))) continue;
            for (uint32_t c = 0; c < L; ++c)
                if (has(fw, fp(a, b, c)) && has(fw, fp(b, a, c)) && has(fw, fp(c, a, b)) && a <= b && b <= c)
                    cs.push_back({fp(a, b, c), {}});
        }
    }
    evaluate(3, cs, nullptr);
    // four-chains v2 - v0 - v1 - v3, :401-459
    cs.clear();
    for (uint32_t a = 0; a < L; ++a) {
        if (!fv[a]) continue;
        for (uint32_t b = 0; b < L; ++b) {
            if (!has(fe, fp(a, b))) continue;
            for (uint32_t c = 0; c < L; ++c) {
                if (!has(fw, fp(a, b, c))) continue;
                for (uint32_t d = 0; d < L; ++d)
                    if (has(fw, fp(b, a, d)) && has(fe, fp(a, c)) && has(fe, fp(b, d)) && a <= b)
                        cs.push_back({fp(a, b, c, d), {}});
            }
        }
    }
    evaluate(4, cs, nullptr);
    // three-stars (leaf, center, leaf, leaf), :461-533
    cs.clear();
    for (uint32_t a = 0; a < L; ++a) {
        if (!fv[a]) continue;
        for (uint32_t b = 0; b < L; ++b) {
            if (!has(fe, fp(a, b))) continue;
            for (uint32_t c = 0; c < L; ++c) {
                if (!has(fw, fp(a, b, c))) continue;
                for (uint32_t d = 0; d < L; ++d)
                    if (has(fw, fp(a, c, d)) && has(fw, fp(a, b, d)) && b <= c && c <= d)
                        cs.push_back({fp(b, a, c, d), {}});
            }
        }
    }
    evaluate(5, cs, nullptr);
    if (ncand) *ncand = nc;
    return 0;
}

}  // extern "C"

