// Plan extraction: turns an app's apply() callback (include/engine.h:589 of
// the reference) into the plan IR of include/kz.h by running it on
// synthetic embeddings.
//
// Why probing: the generated apps (kautomine/*.cc, kgraphpi/*.cc) must link
// unchanged, so the engine only ever sees apply() as a host function
// pointer.  Its behaviour is a fixed sequence of set operations
// (VertexSet::intersect_with / subtracted_by) on neighbour lists and cached
// prefixes, followed by candidate loops that either scatter a child
// (engine->scatter) or add a count (Aggregator::add), filtered by comparisons
// of the candidate with the matched vertices.  Every one of those touches a
// symbol or inline accessor the library provides, so it can be observed.
//
// What is observed, for embeddings of size s and ORDER TYPE t (the relative
// order of the matched vertices, ties included):
//   * each set operation: its kind, the provenance of both operands (N(m_j)
//     buffer, or the result of an earlier operation recorded by pointer) and
//     where its result goes (a cached object or scratch);
//   * each candidate loop: the set it reads (VertexSet::get_vertex /
//     get_vertices_list / trim_v_list* report the list), then either the
//     scattered vertices or the delta added to an aggregator;
//   * the predicate: every set is given the same content X per run, X empty
//     or one vertex in one POSITION CLASS (below / equal to / between /
//     above the matched values).  A count loop's delta (0 or 1) and a
//     scatter loop's calls give one bit of the predicate per class.
// Reaching a type at size s replays the chain from a root: at each ancestor
// level every set holds the chain's later vertices, the probe follows the
// child whose vertex is the next chain vertex, and before the target call
// the cached prefixes the ancestors wrote are overwritten with X (their
// length word sits just before the list, as create_cached_obj lays it out).
// Types reached by a scatter of some class are explored depth-first.
//
// Bounded operations (bounded_intersect_with / bounded_subtracted_by,
// src/vertex_set.cc:113-184: the result keeps only ids below upper_bound)
// become plain operations whose bound lives in the predicates: the bound
// must be a matched vertex m_j, every list derived from a bounded one
// (intersect: either operand; subtract: the minuend) inherits the lowest
// bound, and the probe writes X filtered by that bound into the result, so
// a candidate loop over it learns "v < m_j" as missing predicate bits.  A
// bounded subtrahend has no such form and is refused.
// intersect_and_count (src/vertex_set.cc:88-111) is an intersect into a
// slot of its own followed by a count loop over that slot: the probe logs
// the operation, a read of the slot and returns |X| (filtered by the
// operands' bounds), which apply() passes to Aggregator::add.
//
// Everything runs on the host, single-threaded, once per run(); it costs a
// few thousand apply() calls on lists of at most a few ids.

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <set>
#include <stdexcept>
#include <string>
#include <vector>

#include "distributed_application.h"
#include "engine.h"
#include "kz.h"
#include "probe.h"

namespace kzk {

ProbeHooks* g_hooks = nullptr;

namespace {

constexpr VertexId kBase = 1u << 24;
constexpr VertexId kGap = 1u << 16;
constexpr VertexId kNoBound = ~VertexId(0);

struct Prov {
    int kind = KZ_SET_NONE;  // KZ_SET_NBR / KZ_SET_SLOT
    int a = 0, b = 0;
    bool operator==(const Prov& o) const { return kind == o.kind && a == o.a && b == o.b; }
    bool operator<(const Prov& o) const {
        return kind != o.kind ? kind < o.kind : (a != o.a ? a < o.a : b < o.b);
    }
};

struct Event {
    enum Type { OP, READ, SCATTER, ADD } type;
    int op = 0;
    Prov lhs, rhs, src;
    VertexId v = 0;
    bool active = true, finalized = false;
    int agg = -1;
    uint64_t delta = 0;
};

struct Region {
    VertexId* start;
    size_t cap;  // in VertexIds
    bool cached;
    int level;
};

// Values realising an order type: dense ranks -> kBase + kGap * (rank + 1).
std::vector<VertexId> values_of(const std::vector<int>& ranks) {
    std::vector<VertexId> m;
    for (int r : ranks) m.push_back(kBase + kGap * VertexId(r + 1));
    return m;
}

std::vector<VertexId> distinct_sorted(const std::vector<VertexId>& m) {
    std::vector<VertexId> z(m);
    std::sort(z.begin(), z.end());
    z.erase(std::unique(z.begin(), z.end()), z.end());
    return z;
}

// A vertex in position class c against distinct sorted values z.
VertexId class_value(const std::vector<VertexId>& z, int c) {
    const int q = int(z.size());
    if (c & 1) return z[c / 2];
    const int i = c / 2;
    if (i < q) return z[i] - kGap / 2;
    return z[q - 1] + kGap / 2;
}

// Dense ranks of (m, v).
std::vector<int> ranks_of(const std::vector<VertexId>& m) {
    std::vector<VertexId> z = distinct_sorted(m);
    std::vector<int> r;
    for (VertexId x : m) r.push_back(int(std::lower_bound(z.begin(), z.end(), x) - z.begin()));
    return r;
}

uint32_t code_of(const std::vector<int>& ranks) {
    uint32_t code = 0, pw = 1;
    const uint32_t s = uint32_t(ranks.size());
    for (int r : ranks) {
        code += uint32_t(r) * pw;
        pw *= s;
    }
    return code;
}

// Receiver of the VertexSet operations apply() issues while probed.
// upper_bound: kNoBound, or the bound of a bounded_* operation.
struct OpSink {
    virtual void on_op(int op, const VertexSet* lhs, const VertexSet* rhs, VertexSet* result,
                       VertexId upper_bound) = 0;
    virtual VertexId on_count(const VertexSet* lhs, const VertexSet* rhs) = 0;
    virtual ~OpSink() {}
};

class Prober final : public ProbeHooks, public Khuzdul::AbstractEmbeddingExplorationEngine, public OpSink {
   public:
    Prober(Khuzdul::ApplyFunction fn, int K, const std::vector<void*>& aggs) : fn_(fn), K_(K), aggs_(aggs) {}

    void extract(kz_plan_desc* plan);

    // ---- hooks
    void on_read(const VertexId* list, VertexId n) override {
        (void)n;
        Event e;
        e.type = Event::READ;
        e.src = prov_of(list);
        log_.push_back(e);
    }
    void on_trim(const VertexId* list, VertexId n) override { on_read(list, n); }
    void on_size(const VertexId* list, VertexId n) override {
        if (prov_.count(list)) on_read(list, n);
    }
    void on_add(const void* agg, uint64_t delta) override {
        Event e;
        e.type = Event::ADD;
        e.agg = -1;
        for (size_t i = 0; i < aggs_.size(); ++i)
            if (aggs_[i] == agg) e.agg = int(i);
        if (e.agg < 0) fail("apply() added to an aggregator not allocated by this DistributedApplication");
        e.delta = delta;
        log_.push_back(e);
    }
    void on_create_cached(VertexId* start, VertexId* obj, VertexId size) override {
        // capacity is bounded by the whole cached buffer, registered when the
        // embedding owning it is created
        (void)start;
        (void)obj;
        (void)size;
    }

    // ---- engine view
    void scatter(const VertexId new_v, const bool active, const uint32_t cached_obj_size,
                 Khuzdul::ExtendableEmbedding& parent, const Context ctx, const bool finalized) override {
        (void)parent;
        (void)ctx;
        Event e;
        e.type = Event::SCATTER;
        e.v = new_v;
        e.active = active;
        e.finalized = finalized;
        log_.push_back(e);
        if (cur_level_ == follow_level_ && new_v == follow_v_ && !child_made_) {
            child_made_ = true;
            child_cached_bytes_ = cached_obj_size;
            child_finalized_ = finalized;
        }
    }
    void* alloc_thread_local_scratchpad_buffer(const Context& ctx, const size_t& size) override {
        (void)ctx;
        const size_t words = (size + 3) / 4 + 8;
        scratch_.emplace_back(new VertexId[words]());
        regions_.push_back(Region{scratch_.back().get(), words, false, cur_level_});
        return scratch_.back().get();
    }
    void clear_thread_local_scratchpad_buffer(const Context& ctx) override { (void)ctx; }
    LabelId get_vertex_label(const VertexId v) override {
        (void)v;
        fail("apply() reads vertex labels; labelled mining is not supported by this engine");
        return 0;
    }

    // VertexSet operations while probing (called from the out-of-line
    // VertexSet methods in khuzdul_api.cpp).
    void on_op(int op, const VertexSet* lhs, const VertexSet* rhs, VertexSet* result,
               VertexId upper_bound) override {
        if (upper_bound != kNoBound && std::find(m_cur_.begin(), m_cur_.end(), upper_bound) == m_cur_.end())
            fail("a bounded set operation's upper bound is not a matched vertex");
        const VertexId b = std::min(upper_bound, result_bound(op, lhs, rhs));
        log_op(op, lhs, rhs);
        // write the configured content, cut at the bound, into the result buffer
        const std::vector<VertexId> x = below(content_, b);
        const Region* r = region_of(result->vertices_list_);
        if (!r) fail("a set operation writes outside the scratchpad and cached objects");
        const size_t room = r->cap - size_t(result->vertices_list_ - r->start);
        if (x.size() > room) fail("probe content does not fit a result buffer");
        std::copy(x.begin(), x.end(), result->vertices_list_);
        result->num_vertices_ = VertexId(x.size());
        prov_[result->vertices_list_] = Prov{KZ_SET_SLOT, cur_level_, ops_this_call_ - 1};
        bound_[result->vertices_list_] = b;
        if (r->cached) cached_results_.push_back(CachedResult{result->vertices_list_, VertexId(x.size()), b});
    }
    VertexId on_count(const VertexSet* lhs, const VertexSet* rhs) override {
        const VertexId b = result_bound(KZ_OP_INTERSECT, lhs, rhs);
        log_op(KZ_OP_INTERSECT, lhs, rhs);
        Event e;
        e.type = Event::READ;
        e.src = Prov{KZ_SET_SLOT, cur_level_, ops_this_call_ - 1};
        log_.push_back(e);
        return VertexId(below(content_, b).size());
    }

   private:
    struct CachedResult {
        VertexId* list;
        VertexId len;
        VertexId bound;
    };
    struct OutKey {
        int kind;  // KZ_OUT_*
        Prov src;
        int agg;
        bool operator<(const OutKey& o) const {
            if (kind != o.kind) return kind < o.kind;
            if (!(src == o.src)) return src < o.src;
            return agg < o.agg;
        }
    };
    struct RunLog {
        std::vector<Event> ev;
    };

    [[noreturn]] void fail(const std::string& msg) { throw std::runtime_error(msg); }

    Prov prov_of(const VertexId* p) {
        auto it = prov_.find(p);
        if (it == prov_.end()) fail("apply() used a vertex list the probe cannot attribute");
        return it->second;
    }
    VertexId bound_of(const VertexId* p) const {
        auto it = bound_.find(p);
        return it == bound_.end() ? kNoBound : it->second;
    }
    // Bound a result inherits from its operands.
    VertexId result_bound(int op, const VertexSet* lhs, const VertexSet* rhs) {
        const VertexId rb = bound_of(rhs->vertices_list_);
        if (op == KZ_OP_SUBTRACT && rb != kNoBound)
            fail("apply() subtracts a bounded set; the plan format bounds results, not subtrahends");
        return std::min(bound_of(lhs->vertices_list_), op == KZ_OP_INTERSECT ? rb : kNoBound);
    }
    static std::vector<VertexId> below(const std::vector<VertexId>& x, VertexId b) {
        std::vector<VertexId> y;
        for (VertexId v : x)
            if (v < b) y.push_back(v);
        return y;
    }
    void log_op(int op, const VertexSet* lhs, const VertexSet* rhs) {
        Event e;
        e.type = Event::OP;
        e.op = op;
        e.lhs = prov_of(lhs->vertices_list_);
        e.rhs = prov_of(rhs->vertices_list_);
        ++ops_this_call_;
        log_.push_back(e);
    }
    const Region* region_of(const VertexId* p) const {
        for (auto it = regions_.rbegin(); it != regions_.rend(); ++it)
            if (p >= it->start && p < it->start + it->cap) return &*it;
        return nullptr;
    }

    RunLog run_chain(const std::vector<VertexId>& m, const std::vector<VertexId>& X);
    void apply_at(int level, const std::vector<VertexId>& m, const std::vector<VertexId>& nbr_content,
                  std::vector<std::vector<VertexId>>& cached_bufs);

    Khuzdul::ApplyFunction fn_;
    int K_;
    const std::vector<void*>& aggs_;

    std::map<const VertexId*, Prov> prov_;
    std::map<const VertexId*, VertexId> bound_;  // results of bounded operations (and their descendants)
    std::vector<VertexId> m_cur_;                // matched vertices of the apply() call running
    std::vector<Region> regions_;
    std::vector<std::unique_ptr<VertexId[]>> scratch_;
    std::vector<std::vector<VertexId>> nbr_bufs_;
    std::vector<CachedResult> cached_results_;
    std::vector<Event> log_;
    std::vector<VertexId> content_;
    int cur_level_ = 0;
    int ops_this_call_ = 0;
    int follow_level_ = -1;
    VertexId follow_v_ = 0;
    bool child_made_ = false;
    uint32_t child_cached_bytes_ = 0;
    bool child_finalized_ = false;
};

// Run apply() for the embedding m[0..level) with every N(m_j) and every op
// result holding `content`.
void Prober::apply_at(int level, const std::vector<VertexId>& m, const std::vector<VertexId>& content,
                      std::vector<std::vector<VertexId>>& cached_bufs) {
    Khuzdul::ExtendableEmbedding e;
    std::memset(&e, 0, sizeof(e));
    e.size = level;
    content_ = content;
    m_cur_.assign(m.begin(), m.begin() + level);
    for (int j = 0; j < level; ++j) {
        nbr_bufs_.emplace_back(content.empty() ? std::vector<VertexId>(1, 0) : content);
        VertexId* p = nbr_bufs_.back().data();
        e.matched_vertices[j] = m[j];
        e.matched_vertices_nbrs_ptx[j] = p;
        e.matched_vertices_num_nbrs[j] = VertexId(content.size());
        Prov pv;
        pv.kind = KZ_SET_NBR;
        pv.a = j;
        prov_[p] = pv;
        e.cached_objects[j] = cached_bufs[j].data();
    }
    cur_level_ = level;
    ops_this_call_ = 0;
    Context ctx;
    std::memset(&ctx, 0, sizeof(ctx));
    ctx.depth = level - 1;
    fn_(e, ctx, this);
}

// Replays the chain root -> ... -> m (size s); the target call sees X.
Prober::RunLog Prober::run_chain(const std::vector<VertexId>& m, const std::vector<VertexId>& X) {
    const int s = int(m.size());
    prov_.clear();
    bound_.clear();
    regions_.clear();
    scratch_.clear();
    nbr_bufs_.clear();
    nbr_bufs_.reserve(64);
    cached_results_.clear();
    // cached object buffer of each embedding size (index = size - 1)
    std::vector<std::vector<VertexId>> cached(size_t(s), std::vector<VertexId>(1 << 12, 0));
    for (int j = 0; j < s; ++j) regions_.push_back(Region{cached[j].data(), cached[j].size(), true, j + 1});
    for (int level = 1; level < s; ++level) {
        std::vector<VertexId> A(m.begin() + level, m.end());
        std::sort(A.begin(), A.end());
        A.erase(std::unique(A.begin(), A.end()), A.end());
        log_.clear();
        follow_level_ = level;
        follow_v_ = m[level];
        child_made_ = false;
        apply_at(level, m, A, cached);
        if (!child_made_) fail("probe chain broken: expected vertex was not scattered");
        // child's own cached object: sized as the app asked (+ slack)
        cached[level].assign(std::max<size_t>(1 << 12, child_cached_bytes_ / 4 + 64), 0);
        regions_.push_back(Region{cached[level].data(), cached[level].size(), true, level + 1});
    }
    // overwrite the ancestors' cached prefixes with X (cut at their bounds)
    for (const CachedResult& c : cached_results_) {
        if (c.list[-1] != c.len) fail("cached prefix without its length word before the list");
        const std::vector<VertexId> x = below(X, c.bound);
        std::copy(x.begin(), x.end(), c.list);
        c.list[-1] = VertexId(x.size());
    }
    log_.clear();
    follow_level_ = -1;
    apply_at(s, m, X, cached);
    RunLog r;
    r.ev = log_;
    return r;
}

void Prober::extract(kz_plan_desc* plan) {
    std::memset(plan, 0, sizeof(*plan));
    std::memset(plan->type_table, 0xFF, sizeof(plan->type_table));
    plan->pattern_size = uint32_t(K_);
    plan->num_aggregators = uint32_t(aggs_.size());
    for (size_t i = 0; i < aggs_.size() && i < KZ_MAX_AGGREGATORS; ++i) plan->multiplicity[i] = 1;
    if (aggs_.size() > KZ_MAX_AGGREGATORS) fail("more aggregators than the plan format holds");

    std::vector<std::vector<int>> todo = {{0}};
    std::set<std::pair<int, uint32_t>> seen;
    std::vector<kz_step_list> lists;
    g_hooks = this;
    try {
        while (!todo.empty()) {
            const std::vector<int> ranks = todo.back();
            todo.pop_back();
            const int s = int(ranks.size());
            const uint32_t code = code_of(ranks);
            if (!seen.insert({s, code}).second) continue;
            const std::vector<VertexId> m = values_of(ranks);
            const std::vector<VertexId> z = distinct_sorted(m);
            const int nclass = 2 * int(z.size()) + 1;
            // baseline and one run per class
            std::vector<RunLog> runs;
            runs.push_back(run_chain(m, {}));
            for (int c = 0; c < nclass; ++c) runs.push_back(run_chain(m, {class_value(z, c)}));
            // ops: identical in every run
            std::vector<Event> ops;
            for (const Event& e : runs[0].ev)
                if (e.type == Event::OP) ops.push_back(e);
            for (const RunLog& r : runs) {
                size_t k = 0;
                for (const Event& e : r.ev)
                    if (e.type == Event::OP) {
                        if (k >= ops.size() || e.op != ops[k].op || !(e.lhs == ops[k].lhs) || !(e.rhs == ops[k].rhs))
                            fail("apply() issues different set operations for the same order type");
                        ++k;
                    }
                if (k != ops.size()) fail("apply() issues different set operations for the same order type");
            }
            if (ops.size() > KZ_MAX_OPS) fail("more set operations per embedding than the plan format holds");
            // outputs: scatter loops keyed by their source; count loops by
            // their position in the sequence of adds.
            std::vector<OutKey> outs;
            std::map<OutKey, uint32_t> pred;
            std::vector<int> add_agg;  // from the baseline run
            for (const Event& e : runs[0].ev)
                if (e.type == Event::ADD) add_agg.push_back(e.agg);
            std::vector<Prov> add_src(add_agg.size());
            std::vector<bool> add_src_known(add_agg.size(), false);
            std::vector<uint32_t> add_pred(add_agg.size(), 0);
            for (size_t ri = 0; ri < runs.size(); ++ri) {
                const int c = int(ri) - 1;
                std::vector<Prov> reads;
                size_t ai = 0;
                for (const Event& e : runs[ri].ev) {
                    if (e.type == Event::READ) {
                        reads.push_back(e.src);
                    } else if (e.type == Event::SCATTER) {
                        if (reads.empty()) fail("scatter without a preceding candidate read");
                        OutKey k{KZ_OUT_SCATTER, reads.back(), -1};
                        if (!pred.count(k)) {
                            pred[k] = 0;
                            outs.push_back(k);
                        }
                        if (c >= 0 && e.v == class_value(z, c)) pred[k] |= 1u << c;
                        if (s + 1 >= K_) fail("apply() scatters beyond the pattern size");
                        reads.clear();
                    } else if (e.type == Event::ADD) {
                        if (ai >= add_agg.size() || add_agg[ai] != e.agg)
                            fail("apply() adds to aggregators in a content-dependent order");
                        if (!reads.empty()) {
                            // the count loop's set is the last one read before the add
                            // (every set is non-empty in the per-class runs)
                            Prov src = reads.back();
                            if (add_src_known[ai] && !(add_src[ai] == src))
                                fail("a count loop changes its source set between runs");
                            add_src[ai] = src;
                            add_src_known[ai] = true;
                        }
                        if (ri == 0 && e.delta != 0) fail("a count loop over empty sets added a non-zero count");
                        if (c >= 0) {
                            if (e.delta > 1) fail("a count loop is not additive over candidates");
                            if (e.delta) add_pred[ai] |= 1u << c;
                        }
                        reads.clear();
                        ++ai;
                    }
                }
                if (ai != add_agg.size()) fail("apply() adds a content-dependent number of counts");
            }
            // assemble the step list
            kz_step_list L;
            std::memset(&L, 0, sizeof(L));
            L.num_ops = uint8_t(ops.size());
            for (size_t j = 0; j < ops.size(); ++j) {
                L.ops[j].op = uint8_t(ops[j].op);
                L.ops[j].lhs = kz_set_ref{uint8_t(ops[j].lhs.kind), uint8_t(ops[j].lhs.a), uint8_t(ops[j].lhs.b), 0};
                L.ops[j].rhs = kz_set_ref{uint8_t(ops[j].rhs.kind), uint8_t(ops[j].rhs.a), uint8_t(ops[j].rhs.b), 0};
            }
            int no = 0;
            auto push_out = [&](int kind, const Prov& src, int agg, uint32_t pbits) {
                if (no >= KZ_MAX_OUTS) fail("more candidate loops per embedding than the plan format holds");
                kz_out& o = L.outs[no++];
                o.kind = uint8_t(kind);
                o.aggregator = uint8_t(agg < 0 ? 0 : agg);
                o.child_active = 1;
                o.src = kz_set_ref{uint8_t(src.kind), uint8_t(src.a), uint8_t(src.b), 0};
                o.pred = pbits;
            };
            for (const OutKey& k : outs) {
                push_out(KZ_OUT_SCATTER, k.src, -1, pred[k]);
                for (int c = 0; c < nclass; ++c) {
                    if (!(pred[k] & (1u << c))) continue;
                    std::vector<VertexId> mc(m);
                    mc.push_back(class_value(z, c));
                    todo.push_back(ranks_of(mc));
                }
            }
            for (size_t i = 0; i < add_agg.size(); ++i) {
                if (!add_src_known[i]) {
                    if (add_pred[i]) fail("a count loop with no attributable source");
                    continue;  // never counts anything
                }
                push_out(KZ_OUT_COUNT, add_src[i], add_agg[i], add_pred[i]);
            }
            L.num_outs = uint8_t(no);
            int idx = -1;
            for (size_t i = 0; i < lists.size(); ++i)
                if (std::memcmp(&lists[i], &L, sizeof(L)) == 0) idx = int(i);
            if (idx < 0) {
                if (lists.size() >= KZ_MAX_STEP_LISTS) fail("too many distinct step lists");
                lists.push_back(L);
                idx = int(lists.size()) - 1;
            }
            plan->type_table[s][code] = uint8_t(idx);
        }
    } catch (...) {
        g_hooks = nullptr;
        throw;
    }
    g_hooks = nullptr;
    plan->num_step_lists = uint32_t(lists.size());
    for (size_t i = 0; i < lists.size(); ++i) plan->lists[i] = lists[i];
}

// ------------------------------------------------------------ labelled apps
// A labelled app (kautomine/three_fsm.cc) runs once per candidate pattern:
// its apply() (:52-250) follows global pattern labels, extends an embedding
// by neighbours whose label matches the next pattern vertex, and at the
// last one inserts every matched vertex into a DistributedPatternDomain
// with `enable` = the candidate's label matches (and it repeats no matched
// vertex).  The probe recovers that pattern: at each size s it runs apply()
// with every neighbour list holding one fresh vertex x, once per label of
// x; exactly one label must lead to a scatter of x or to an enabled insert
// of x.  The list the candidate loop read gives x's neighbours in the
// pattern (N(m_j): j; an intersection: every operand's vertex).  A run with
// the list holding a matched vertex m_j that could also be x (same label,
// adjacent to all of x's neighbours) must be rejected.  Inserts end the
// pattern: every position is inserted with its matched vertex.
class LabelProber final : public ProbeHooks, public Khuzdul::AbstractEmbeddingExplorationEngine, public OpSink {
   public:
    LabelProber(Khuzdul::ApplyFunction fn, int K, LabelId L) : fn_(fn), K_(K), L_(L) {}

    void extract(bool root_known, LabelId root_label, LabeledPlan* out);

    // ---- hooks
    void on_read(const VertexId* list, VertexId n) override {
        (void)n;
        Ev e;
        e.type = Ev::READ;
        e.src = prov_of(list);
        log_.push_back(e);
    }
    void on_trim(const VertexId* list, VertexId n) override { on_read(list, n); }
    void on_add(const void*, uint64_t) override { fail("a labelled apply() adds to an aggregator"); }
    void on_create_cached(VertexId*, VertexId*, VertexId) override {}
    void on_domain_insert(const void* domain, int pattern_vertex, VertexId v, bool enable) override {
        Ev e;
        e.type = Ev::INSERT;
        e.domain = domain;
        e.idx = pattern_vertex;
        e.v = v;
        e.enable = enable;
        log_.push_back(e);
    }
    void scatter(const VertexId new_v, const bool, const uint32_t, Khuzdul::ExtendableEmbedding&, const Context,
                 const bool) override {
        Ev e;
        e.type = Ev::SCATTER;
        e.v = new_v;
        log_.push_back(e);
    }
    void* alloc_thread_local_scratchpad_buffer(const Context&, const size_t& size) override {
        const size_t words = (size + 3) / 4 + 8;
        scratch_.emplace_back(new VertexId[words]());
        regions_.push_back(Region{scratch_.back().get(), words, false, 0});
        return scratch_.back().get();
    }
    void clear_thread_local_scratchpad_buffer(const Context&) override {}
    LabelId get_vertex_label(const VertexId v) override {
        auto it = label_.find(v);
        if (it == label_.end()) fail("apply() reads the label of a vertex the probe did not place");
        return it->second;
    }
    VertexId on_count(const VertexSet*, const VertexSet*) override {
        fail("a labelled apply() counts an intersection (its pattern inserts each matched vertex)");
        return 0;
    }
    void on_op(int op, const VertexSet* lhs, const VertexSet* rhs, VertexSet* result,
               VertexId upper_bound) override {
        if (op != KZ_OP_INTERSECT) fail("a labelled apply() subtracts sets (induced patterns are not supported)");
        if (upper_bound != kNoBound) fail("a labelled apply() bounds a set operation");
        Ev e;
        e.type = Ev::OP;
        e.lhs = prov_of(lhs->vertices_list_);
        e.rhs = prov_of(rhs->vertices_list_);
        const int idx = int(ops_.size());
        ops_.push_back(e);
        log_.push_back(e);
        const Region* r = nullptr;
        for (auto it = regions_.rbegin(); it != regions_.rend() && !r; ++it)
            if (result->vertices_list_ >= it->start && result->vertices_list_ < it->start + it->cap) r = &*it;
        if (!r) fail("a set operation writes outside the scratchpad");
        // every list holds the same content, so the intersection is that content
        if (content_.size() > r->cap - size_t(result->vertices_list_ - r->start))
            fail("probe content does not fit a result buffer");
        std::copy(content_.begin(), content_.end(), result->vertices_list_);
        result->num_vertices_ = VertexId(content_.size());
        Prov pv;
        pv.kind = KZ_SET_SLOT;
        pv.b = idx;
        prov_[result->vertices_list_] = pv;
    }

   private:
    struct Ev {
        enum Type { READ, OP, SCATTER, INSERT } type;
        Prov src, lhs, rhs;
        const void* domain = nullptr;
        int idx = 0;
        VertexId v = 0;
        bool enable = false;
    };
    [[noreturn]] void fail(const std::string& msg) { throw std::runtime_error(msg); }
    Prov prov_of(const VertexId* p) {
        auto it = prov_.find(p);
        if (it == prov_.end()) fail("apply() used a vertex list the probe cannot attribute");
        return it->second;
    }
    // Pattern vertices whose neighbour lists feed the set `p`.
    uint32_t sources(const Prov& p) {
        if (p.kind == KZ_SET_NBR) return 1u << p.a;
        if (p.kind == KZ_SET_SLOT && p.b < int(ops_.size())) return sources(ops_[p.b].lhs) | sources(ops_[p.b].rhs);
        fail("a candidate loop reads a set the probe cannot attribute");
    }
    std::vector<Ev> run(const std::vector<VertexId>& m, const std::vector<VertexId>& content) {
        const int s = int(m.size());
        prov_.clear();
        regions_.clear();
        scratch_.clear();
        ops_.clear();
        log_.clear();
        nbr_bufs_.assign(size_t(s), content);
        cached_.assign(size_t(s), std::vector<VertexId>(1 << 12, 0));
        content_ = content;
        Khuzdul::ExtendableEmbedding e;
        std::memset(&e, 0, sizeof(e));
        e.size = s;
        for (int j = 0; j < s; ++j) {
            e.matched_vertices[j] = m[j];
            e.matched_vertices_nbrs_ptx[j] = nbr_bufs_[j].data();
            e.matched_vertices_num_nbrs[j] = VertexId(content.size());
            e.cached_objects[j] = cached_[j].data();
            regions_.push_back(Region{cached_[j].data(), cached_[j].size(), true, j + 1});
            Prov pv;
            pv.kind = KZ_SET_NBR;
            pv.a = j;
            prov_[nbr_bufs_[j].data()] = pv;
        }
        Context ctx;
        std::memset(&ctx, 0, sizeof(ctx));
        ctx.depth = s - 1;
        fn_(e, ctx, this);
        return log_;
    }

    Khuzdul::ApplyFunction fn_;
    int K_;
    LabelId L_;
    std::map<VertexId, LabelId> label_;
    std::map<const VertexId*, Prov> prov_;
    std::vector<Region> regions_;
    std::vector<std::unique_ptr<VertexId[]>> scratch_;
    std::vector<std::vector<VertexId>> nbr_bufs_, cached_;
    std::vector<Ev> ops_, log_;
    std::vector<VertexId> content_;
};

void LabelProber::extract(bool root_known, LabelId root_label, LabeledPlan* out) {
    *out = LabeledPlan();
    auto vid = [](int i) { return kBase + kGap * VertexId(i + 1); };
    std::vector<VertexId> m = {vid(0)};
    std::vector<LabelId> lab = {root_label};
    std::vector<uint32_t> adj = {0};
    if (!root_known) {
        // the root's label: the one for which apply() does anything
        int found = -1;
        for (LabelId l = 0; l < L_; ++l) {
            bool any = false;
            for (LabelId c = 0; c < L_ && !any; ++c) {
                label_.clear();
                label_[m[0]] = l;
                label_[vid(1)] = c;
                const std::vector<Ev> ev = run(m, {vid(1)});
                for (const Ev& e : ev) any |= e.type != Ev::READ;
            }
            if (any) {
                if (found >= 0) fail("apply() accepts roots of more than one label: run() needs the starting label");
                found = int(l);
            }
        }
        if (found < 0) {
            out->empty = true;
            return;
        }
        lab[0] = LabelId(found);
    }
    for (int s = 1;; ++s) {
        if (s >= MAX_PATTERN_SIZE) fail("a labelled pattern larger than MAX_PATTERN_SIZE");
        const VertexId x = vid(s);
        label_.clear();
        for (int j = 0; j < s; ++j) label_[m[j]] = lab[j];
        int acc = -1;
        bool inserts = false, scatters = false;
        uint32_t srcs = 0;
        std::vector<Ev> acc_ev;
        for (LabelId l = 0; l < L_; ++l) {
            label_[x] = l;
            const std::vector<Ev> ev = run(m, {x});
            bool takes = false;
            Prov last_read;
            bool have_read = false;
            for (const Ev& e : ev) {
                if (e.type == Ev::READ) {
                    last_read = e.src;
                    have_read = true;
                } else if (e.type == Ev::SCATTER && e.v == x) {
                    scatters = true;
                    if (!have_read) fail("a scatter without a preceding candidate read");
                    takes = true;
                    srcs = sources(last_read);
                } else if (e.type == Ev::INSERT) {
                    inserts = true;
                    if (e.idx == s && e.v == x && e.enable) {
                        if (!have_read) fail("an insert without a preceding candidate read");
                        takes = true;
                        srcs = sources(last_read);
                    }
                }
            }
            if (takes) {
                if (acc >= 0) fail("apply() takes a candidate of more than one label at one pattern vertex");
                acc = int(l);
                acc_ev = ev;
            }
        }
        if (inserts && scatters) fail("apply() both scatters and inserts at one size");
        if (acc < 0) {
            out->empty = true;  // no label extends the pattern here: no embedding
            return;
        }
        // a matched vertex that could also be the candidate must be refused
        for (int j = 0; j < s; ++j) {
            if (lab[j] != LabelId(acc) || (srcs & ~adj[j]) != 0) continue;
            const std::vector<Ev> ev = run(m, {m[j]});
            for (const Ev& e : ev)
                if ((e.type == Ev::SCATTER && e.v == m[j]) ||
                    (e.type == Ev::INSERT && e.idx == s && e.v == m[j] && e.enable))
                    fail("apply() lets one vertex take two pattern vertices");
        }
        lab.push_back(LabelId(acc));
        adj.push_back(srcs);
        for (int j = 0; j < s; ++j)
            if (srcs & (1u << j)) adj[j] |= 1u << s;
        if (!inserts) {
            if (s + 1 >= K_) fail("apply() scatters beyond the pattern size");
            m.push_back(x);
            continue;
        }
        // the pattern ends here: every position inserted with its vertex
        const int k = s + 1;
        std::vector<bool> seen(size_t(k), false);
        for (const Ev& e : acc_ev) {
            if (e.type != Ev::INSERT) continue;
            if (e.idx < 0 || e.idx >= k) fail("apply() inserts at a pattern vertex it did not match");
            const VertexId want = e.idx < s ? m[e.idx] : x;
            if (e.v != want || !e.enable) fail("apply() inserts a vertex other than the matched one");
            if (out->domain && out->domain != e.domain) fail("apply() inserts into two pattern domains");
            out->domain = const_cast<void*>(e.domain);
            seen[size_t(e.idx)] = true;
        }
        for (int i = 0; i < k; ++i)
            if (!seen[size_t(i)]) fail("apply() leaves a pattern vertex out of its domain");
        out->size = k;
        for (int i = 0; i < k; ++i) {
            out->labels[i] = lab[size_t(i)];
            out->adj[i] = adj[size_t(i)];
        }
        return;
    }
}

OpSink* g_prober = nullptr;

}  // namespace

static void require_prober() {
    if (!g_prober) {
        std::fprintf(stderr,
                     "VertexSet set operations run on the GPU inside DistributedApplication::run(); "
                     "they cannot be called from host code outside apply().\n");
        std::exit(-1);
    }
}

void probe_set_op(int op, const VertexSet* lhs, const VertexSet* rhs, VertexSet* result, VertexId upper_bound) {
    require_prober();
    g_prober->on_op(op, lhs, rhs, result, upper_bound);
}

void probe_set_op(int op, const VertexSet* lhs, const VertexSet* rhs, VertexSet* result) {
    probe_set_op(op, lhs, rhs, result, kNoBound);
}

VertexId probe_count_op(const VertexSet* lhs, const VertexSet* rhs) {
    require_prober();
    return g_prober->on_count(lhs, rhs);
}

int extract_labeled_pattern(Khuzdul::ApplyFunction fn, int max_pattern_size, LabelId num_labels, bool root_known,
                            LabelId root_label, LabeledPlan* out, std::string* error) {
    LabelProber p(fn, max_pattern_size, num_labels);
    g_prober = &p;
    g_hooks = &p;
    int rc = 0;
    try {
        p.extract(root_known, root_label, out);
    } catch (const std::exception& ex) {
        if (error) *error = ex.what();
        rc = -1;
    }
    g_hooks = nullptr;
    g_prober = nullptr;
    return rc;
}

int extract_plan(Khuzdul::ApplyFunction fn, int max_pattern_size, const std::vector<void*>& aggregators,
                 kz_plan_desc* plan, std::string* error) {
    Prober p(fn, max_pattern_size, aggregators);
    g_prober = &p;
    int rc = 0;
    try {
        p.extract(plan);
    } catch (const std::exception& ex) {
        if (error) *error = ex.what();
        rc = -1;
    }
    g_prober = nullptr;
    return rc;
}

}  // namespace kzk
