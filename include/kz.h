/*
 * kz.h — C ABI of the MI355X embedding-extension engine (libkz_hip.so).
 *
 * This is the drop-in boundary between host code (the Khuzdul application
 * API in openkhuzdulgraph_amd/khuzdul, the Python host mirror, or any FFI
 * caller) and the gfx950 HIP kernels.  Signatures use plain pointers, sizes
 * and integer handles only.  Every function returns an int: 0 on success, a
 * negative KZ_ERR_* code on failure; kz_last_error() then returns a message
 * for the calling thread.  No C++ exception crosses this boundary.
 *
 * Ownership: host buffers passed in are caller-owned and only read during
 * the call; kz_csr arrays returned by the library are released with
 * kz_csr_release(); device memory behind kz_graph / kz_plan handles is
 * library-owned and released with the matching *_release call.
 *
 * Each declaration names the reference interface it replaces (paths relative
 * to the OpenKhuzdulGraph source tree).
 */
#ifndef KZ_H
#define KZ_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define KZ_OK 0
#define KZ_ERR_INVALID (-1)     /* bad argument or malformed input           */
#define KZ_ERR_NOMEM (-2)       /* host or device allocation failed          */
#define KZ_ERR_IO (-3)          /* file could not be read or written         */
#define KZ_ERR_DEVICE (-4)      /* HIP runtime error / no usable device      */
#define KZ_ERR_UNSUPPORTED (-5) /* plan or pattern outside what is supported */
#define KZ_ERR_BUDGET (-6)      /* request exceeds the configured HBM budget */
#define KZ_ERR_STATE (-7)       /* call made in the wrong state (e.g. no init) */

/* ------------------------------------------------------------------------ */
/* Host-side CSR graph (row a10: graph storage + .dgraph loader)            */
/* ------------------------------------------------------------------------ */

/* Compressed sparse rows with 64-bit row offsets and 32-bit column ids.
 * Replaces HashBasedNumaAwareDistributedCSRGraph (include/graph.h:308-425)
 * on one device: the whole graph, not a hash partition.  Neighbour lists are
 * strictly increasing.  `oriented` is 1 when every stored edge u->v has
 * v > u (degree-ordered rank ids, toolkits/to_oriented_dataset_converter.cc). */
typedef struct kz_csr {
    uint32_t num_vertices;
    uint32_t max_degree;
    uint64_t num_entries; /* stored directed entries = .dgraph num_edges   */
    uint32_t oriented;
    uint32_t has_labels;
    uint64_t* rowptr; /* num_vertices + 1                                 */
    uint32_t* col;    /* num_entries                                      */
    uint32_t* labels; /* num_vertices, or NULL                            */
} kz_csr;

/* Read a .dgraph file: packed 20-byte header {u32 nv; u64 ne; i32
 * has_node_label; i32 has_edge_label}, optional u32 labels[nv], then ne
 * packed {u32 src, u32 dst} sorted by (src, dst).
 * Replaces JointHashBasedNumaAwareDistGraphLoader::load_graphs
 * (include/graph_loader.h:706-961) and its format (graph_loader.h:31-42).
 * `oriented` is detected (every dst > src).  Edge labels are rejected, as in
 * the reference (graph_loader.h:770). */
int kz_dgraph_load(const char* path, kz_csr* out);

/* Write a CSR in the same format (toolkits/converter.cc:178-263 for
 * symmetric graphs, to_oriented_dataset_converter.cc:66-117 for oriented). */
int kz_dgraph_store(const char* path, const kz_csr* g);

/* Free arrays owned by a kz_csr returned by this library; zeroes *g. */
void kz_csr_release(kz_csr* g);

/* Build a clean symmetric CSR from an undirected edge list: drop self-loops
 * and duplicates, drop isolated vertices and relabel the rest 0..n'-1 in
 * ascending original id (datasets/live_journal/standardize.cc:66-114), store
 * both directions sorted (toolkits/converter.cc:237-263).  Ids must be
 * < num_vertices_hint. */
int kz_csr_from_edges(const uint32_t* src, const uint32_t* dst, uint64_t num_edges,
                      uint32_t num_vertices_hint, kz_csr* out);

/* Seeded synthetic graphs, cleaned as kz_csr_from_edges.  All randomness is
 * a counter-based splitmix64 stream, so the bytes are identical on every
 * host and thread count. */
int kz_gen_erdos_renyi(uint32_t n, uint64_t m, uint64_t seed, kz_csr* out);
int kz_gen_rmat(uint32_t scale, uint32_t edge_factor, double a, double b, double c,
                uint64_t seed, kz_csr* out);
int kz_gen_chung_lu(uint32_t n, uint64_t num_draws, double exponent, uint32_t max_weight,
                    uint64_t seed, kz_csr* out);

/* Degree orientation: rank vertices by (degree asc, id asc), relabel by rank,
 * keep u->v iff rank(v) > rank(u) (toolkits/to_oriented_dataset_converter.cc:
 * 27-124).  `sym` must be symmetric. */
int kz_orient_by_degree(const kz_csr* sym, kz_csr* out);


/* Structural checks of toolkits/check_graph.cc:44-68: strictly increasing
 * lists, no self-loops, degree sum = entries; symmetric graphs also need
 * every reverse edge, oriented graphs every dst > src.  Returns KZ_OK or
 * KZ_ERR_INVALID with a message. */
int kz_csr_check(const kz_csr* g);

/* 64-bit digest (splitmix64 fold of n, entries, rowptr[], col[]; thread-count
 * independent) used to prove that two builds produced the same graph. */
uint64_t kz_csr_digest(const kz_csr* g);

/* Attach vertex labels (labels[num_vertices], copied; replaces any held)
 * to a host CSR: .dgraph files store them after the header
 * (graph_loader.h:248-257) and kz_graph_upload uploads them.  Labels are
 * LabelId (u32) values below KZ_MAX_LABELS (MAX_NUM_LABELS,
 * graph_loader.h:28). */
#define KZ_MAX_LABELS 48
int kz_csr_set_labels(kz_csr* g, const uint32_t* labels);

/* ------------------------------------------------------------------------ */
/* Device runtime                                                            */
/* ------------------------------------------------------------------------ */

/* Select the HIP device and create the library's stream.  Replaces
 * DistributedSys::init_distributed_sys (src/distributed_sys.cc:23-26) for the
 * device side.  Safe to call again with the same device. */
int kz_init(int device);
int kz_shutdown(void);
const char* kz_last_error(void);

typedef struct kz_device_info {
    char name[64];
    char arch[32];
    int compute_units;
    int warp_size;
    uint64_t total_mem;
    uint64_t lds_per_block;
    int clock_khz;
    int mem_clock_khz;
    int mem_bus_width;
    int l2_bytes;
} kz_device_info;
int kz_get_device_info(kz_device_info* out);

/* Opaque handle to a CSR resident in HBM. */
typedef struct kz_graph_s* kz_graph;

/* Copy a host CSR into HBM (row offsets u64, columns u32, out-degree u32).
 * Replaces the per-socket CSR build of graph_loader.h:873-901. */
int kz_graph_upload(const kz_csr* g, kz_graph* out);
int kz_graph_release(kz_graph g);
int kz_graph_num_vertices(kz_graph g, uint32_t* out);
/* Vertices, stored entries, largest out-degree and whether it is oriented. */
int kz_graph_info(kz_graph g, uint32_t* num_vertices, uint64_t* num_entries, uint32_t* max_degree, int* oriented);

/* The same generators on the GPU, building the graph directly in HBM:
 * identical draws and the same CSR bit for bit as kz_gen_* (+
 * kz_orient_by_degree when orient != 0), with the clean-up (self-loops,
 * duplicates, isolated vertices, ascending relabel;
 * datasets/live_journal/standardize.cc:66-114) and the orientation
 * (toolkits/to_oriented_dataset_converter.cc:27-124) done by radix sorts.
 * Needs about 24 bytes of device memory per draw while it runs. */
int kz_gpu_gen_erdos_renyi(uint32_t n, uint64_t m, uint64_t seed, int orient, kz_graph* out);
int kz_gpu_gen_rmat(uint32_t scale, uint32_t edge_factor, double a, double b, double c, uint64_t seed, int orient,
                    kz_graph* out);
int kz_gpu_gen_chung_lu(uint32_t n, uint64_t num_draws, double exponent, uint32_t max_weight, uint64_t seed,
                        int orient, kz_graph* out);
/* Copies a resident graph back into a host CSR (released with
 * kz_csr_release); e.g. to write it as .dgraph. */
int kz_graph_download(kz_graph g, kz_csr* out);
/* Degree orientation of a resident symmetric graph on the GPU: a new
 * resident graph, the CSR kz_orient_by_degree builds on the host (rank by
 * (degree asc, id asc), relabel by rank, keep u->v iff rank(v) > rank(u),
 * lists sorted).  Replaces toolkits/to_oriented_dataset_converter.cc:27-124,
 * the offline step the clique apps' oriented inputs come from.  The clique
 * counts of a symmetric graph run on this orientation of it. */
int kz_graph_orient(kz_graph sym, kz_graph* out);
/* Largest number of keys one radix-sort call of the GPU generators takes
 * (default 2^30; smaller values exercise the chunked path in tests). */
int kz_set_sort_chunk(uint64_t items);

/* ------------------------------------------------------------------------ */
/* Counting                                                                  */
/* ------------------------------------------------------------------------ */

#define KZ_MAX_PATTERN 5     /* MAX_PATTERN_SIZE (include/engine.h:23)        */
#define KZ_MAX_AGGREGATORS 8
#define KZ_MAX_OPS 4         /* set operations per (size, order type)         */
#define KZ_MAX_OUTS 4        /* scatter / count outputs per (size, order type) */
#define KZ_MAX_STEP_LISTS 128

#define KZ_METHOD_PLAIN 0      /* level-by-level: materialised prefixes and
                                  embedding queues, chunked by an HBM budget
                                  (src/engine.cc:914-1118)                    */
#define KZ_METHOD_PERSISTENT 1 /* degree-bucketed persistent kernels over a
                                  device work queue, one hipGraph launch      */

typedef struct kz_run_stats {
    double device_ms;                   /* HIP-event time on the library stream */
    double wall_ms;                     /* host wall time of the call           */
    uint64_t launches;                  /* kernel or graph launches             */
    uint64_t tasks;                     /* root tasks (persistent form)         */
    uint64_t embeddings[KZ_MAX_PATTERN + 1]; /* embeddings extended per size    */
    uint64_t chunks;                    /* HBM-budget chunks (plain form)       */
    uint64_t peak_bytes;                /* peak device scratch of the call      */
} kz_run_stats;

/* Exact k-clique count (k = 3..5).  Equals the count printed by
 * kautomine/{triangle,four_clique,five_clique}.cc and the kgraphpi apps of
 * the same names on the same .dgraph, symmetric or degree-oriented.  The
 * persistent form runs on the degree orientation of a symmetric graph
 * (kz_graph_orient, built once and kept with the graph); a graph whose
 * forward out-degree still exceeds 8192 is counted by the plain form. */
int kz_count_cliques(kz_graph g, int k, int method, uint64_t* count, kz_run_stats* stats);

/* Degree buckets of the persistent form: root tasks with |N+(u)| <= bound[0]
 * run one per wave (bound[0] <= 64); up to bound[1..3] one per workgroup
 * with bitmap rows in LDS (bound[3] <= 1024); larger ones keep rows in HBM.
 * block[b] = threads per workgroup of bucket b (256, 512 or 1024; bucket 0:
 * 256).  serial != 0 runs the bucket kernels one after another, each with
 * the whole GPU, instead of as concurrent branches of the captured graph.
 * No reference counterpart: the reference balances roots with 64-embedding
 * chunks per thread (src/engine.cc:985-989). */
typedef struct kz_bucket_cfg {
    uint32_t bound[4];
    uint32_t block[5];
    uint32_t serial;
} kz_bucket_cfg;
int kz_clique_get_buckets(kz_graph g, int k, kz_bucket_cfg* out);
int kz_clique_set_buckets(kz_graph g, int k, const kz_bucket_cfg* cfg);
/* Autotuner: times every candidate configuration (reps runs each, after one
 * warm-up) on this graph and k, keeps the fastest for later calls and
 * returns it with its average device time. */
int kz_clique_tune(kz_graph g, int k, uint32_t reps, kz_bucket_cfg* chosen, double* best_ms);

/* Batched form: many (graph, k) queries in one captured hipGraph launch.
 * The batch keeps the graphs' handles, which must outlive it.  Query
 * (g, j) counts ks[j]-cliques of graphs[g]; kz_batch_run writes
 * counts[g * num_ks + j]. */
typedef struct kz_batch_s* kz_batch;
int kz_batch_create(const kz_graph* graphs, uint32_t num_graphs, const int* ks, uint32_t num_ks, kz_batch* out);
int kz_batch_run(kz_batch b, uint64_t* counts, kz_run_stats* stats);
int kz_batch_tune(kz_batch b, uint32_t reps, kz_bucket_cfg* chosen, double* best_ms);
int kz_batch_release(kz_batch b);

/* Exact terms of the merge-model byte count of the clique schedules on the
 * forward (v > u) graph, for throughput reporting:
 *   out[0] = m              out[1] = T = sum_{(u,v)} c(u,v)
 *   out[2] = sum c(u,v)^2   out[3] = sum_{(u,v,w) in triangles} d+(w)
 *   out[4] = sum_u d+(u)^2  out[5] = sum_{(u,v)} d+(v)
 * with c(u,v) = |N+(u) & N+(v)|.  Then B3 = 4(m + out[4] + out[5]) and
 * B4 = B3 + 4(T + out[2] + out[3]) bytes.  No reference counterpart: the
 * reference reports no bandwidth for this path. */
int kz_clique_workload(kz_graph g, uint64_t* out6);

/* ---- Graphs beyond HBM ---------------------------------------------------
 * A degree-oriented CSR whose neighbour lists stay in pinned host memory
 * (hipHostMalloc; the row offsets and degrees, about 36 bytes per vertex
 * with the working arrays, are in HBM), counted in batches of consecutive
 * roots.  Per batch the lists the batch needs -- its roots' and their
 * forward neighbours' -- are requested once per vertex (a device-side
 * deduplicated request list), served from an HBM cache of the
 * highest-degree vertices' lists or read from host memory over PCIe into a
 * working CSR of list_bytes, and the batch's roots are counted there by the
 * persistent clique kernels.  Replaces the remote graph-data fetch of the
 * distributed engine: GraphDataRequestSender / its handler with the cache of
 * high-degree lists (src/engine.cc:94-510, :513-647; cache_degree_threshold
 * and relative_cache_size of DistributedApplication,
 * include/distributed_application.h:38-46).  A root whose own lists exceed
 * list_bytes fails kz_paged_create with KZ_ERR_BUDGET. */
typedef struct kz_paged_s* kz_paged;
typedef struct kz_paged_cfg {
    uint64_t list_bytes;              /* HBM for the lists of one root batch  */
    uint64_t cache_bytes;             /* HBM for the cache of lists           */
    uint32_t cache_degree_threshold;  /* total degree for a list to be cached */
    uint32_t reserved;
} kz_paged_cfg;
typedef struct kz_paged_stats {
    double device_ms;          /* HIP-event time of all batches          */
    double wall_ms;
    uint64_t batches;
    uint64_t requests;         /* distinct vertices fetched, all batches */
    uint64_t request_refs;     /* list references before deduplication   */
    uint64_t host_bytes;       /* list bytes read from host memory       */
    uint64_t cache_bytes;      /* list bytes served by the HBM cache     */
    uint64_t cached_vertices;  /* lists the cache holds                  */
} kz_paged_stats;
int kz_paged_create(const kz_csr* oriented, const kz_paged_cfg* cfg, kz_paged* out);
/* Exact k-clique count (k = 3..5), as kz_count_cliques. */
int kz_paged_count_cliques(kz_paged p, int k, uint64_t* count, kz_paged_stats* stats);
int kz_paged_release(kz_paged p);

/* ---- Motif census -------------------------------------------------------
 * Induced connected 3- and 4-vertex subgraph counts of a resident SYMMETRIC
 * graph, counts[0..8) = {wedge (open 2-path), triangle, 3-path, 3-star,
 * 4-cycle, tailed triangle, diamond, 4-clique}.  counts[0..2) are what the
 * three_motif apps print (kautomine/three_motif.cc:59-139 with the
 * multiplicities of :44; the kgraphpi variant counts non-induced wedges,
 * kgraphpi/three_motif.cc:118-162); the reference has no 4-vertex census
 * app.  The first call relabels the graph by (degree, id) rank on the GPU
 * and keeps that copy with the graph.  Oriented graphs are rejected
 * (KZ_ERR_INVALID). */
#define KZ_CENSUS_PATTERNS 8
int kz_motif_census(kz_graph g, uint64_t* counts, kz_run_stats* stats);

/* ---------------------------------------------- labelled pattern mining -- */

/* Vertex labels of a resident symmetric graph (labels[num_vertices], each
 * below KZ_MAX_LABELS), uploaded with the label-sorted neighbour lists that
 * frequent subgraph mining reads.  kz_graph_upload calls it for a kz_csr
 * that carries labels.  Replaces the labelled load of
 * DistributedApplication(..., load_vertex_labels = true)
 * (distributed_application.h:38-46, graph_loader.h:248-257, :379-420). */
int kz_graph_set_labels(kz_graph g, const uint32_t* labels);

/* Levels of kz_fsm_mine, in the order kautomine/three_fsm.cc mines them. */
#define KZ_FSM_VERTEX 0
#define KZ_FSM_EDGE 1
#define KZ_FSM_WEDGE 2       /* labels: center, leaf, leaf               */
#define KZ_FSM_TRIANGLE 3
#define KZ_FSM_FOUR_CHAIN 4  /* labels of the path v2 - v0 - v1 - v3: v0, v1, v2, v3 */
#define KZ_FSM_THREE_STAR 5  /* labels: leaf, center, leaf, leaf        */
#define KZ_FSM_LEVELS 6

typedef struct kz_fsm_stats {
    uint32_t num_frequent[KZ_FSM_LEVELS]; /* "Number of frequent ... patterns" */
    uint64_t support_sum[KZ_FSM_LEVELS];  /* "support sum"; level 0 sums every label */
    uint32_t num_candidates[KZ_FSM_LEVELS];
    uint32_t num_labels;
    double device_ms; /* domain kernels and popcounts            */
    double total_ms;  /* including candidate generation         */
} kz_fsm_stats;

typedef struct kz_fsm_candidate {
    uint32_t level;     /* KZ_FSM_EDGE .. KZ_FSM_THREE_STAR          */
    uint32_t labels[4]; /* canonical labels (unused positions 0)     */
    uint32_t support;   /* MNI: smallest per-position domain         */
    uint32_t frequent;  /* support >= threshold                      */
} kz_fsm_candidate;

/* Frequent labelled 1..4-vertex patterns with minimum-image support >=
 * threshold, level by level as kautomine/three_fsm.cc
 * mining_frequent_patterns (:688-695) with its candidate generators
 * (:276-604) and DistributedPatternDomain::get_support (engine.h:476-560).
 * Every evaluated candidate is written to cands (up to cap, NULL allowed);
 * *ncand (may be NULL) gets their number.  Needs kz_graph_set_labels.
 * The reference prints the stats' num_frequent / support_sum per level;
 * it has no other output. */
int kz_fsm_mine(kz_graph g, uint32_t threshold, kz_fsm_stats* stats, kz_fsm_candidate* cands, uint32_t cap,
                uint32_t* ncand);

/* The minimum-image domains of ONE labelled pattern of a kz_fsm_mine level
 * (KZ_FSM_EDGE .. KZ_FSM_THREE_STAR; labels in that level's position order,
 * see the KZ_FSM_* comments): sizes[i] = how many distinct vertices take
 * position i in some embedding (sizes of unused positions are 0).  This is
 * what one run() of kautomine/three_fsm.cc collects in its
 * DistributedPatternDomain (engine.h:405-580) through insert() calls from
 * apply() (three_fsm.cc:52-250), the support being the smallest size
 * (get_support, engine.h:486-560); libkhuzdul_amd.so calls it for the
 * unchanged app.  Needs kz_graph_set_labels. */
int kz_fsm_domains(kz_graph g, uint32_t level, const uint32_t* labels, uint32_t* sizes);

/* ---- Serving --------------------------------------------------------------
 * A server runs queries on resident graphs from a worker thread.  Queries
 * are admitted in arrival order while the device memory of the prepared
 * executions the server keeps (root task lists, bitmap slabs, census
 * copies) stays within hbm_budget bytes; idle prepared executions are
 * released, least recently used first, to make room, and a query whose own
 * execution exceeds the budget fails with KZ_ERR_BUDGET.  Admitted queries
 * are launched as soon as they are prepared, up to max_inflight at once on
 * separate streams, so the device does not idle while queries wait
 * (continuous batching); two queries that need the same prepared execution
 * run one after the other.  A prepared execution already kept by a graph
 * (from a direct call) is adopted and accounted when a query needs it.  Until
 * kz_server_stop returns, the graphs submitted to a server must stay alive
 * and must not be used by other calls.  No reference
 * counterpart: the reference runs one app per process
 * (src/distributed_application.cc:259-356). */
typedef struct kz_server_s* kz_server;
typedef struct kz_server_cfg {
    uint64_t hbm_budget;    /* bytes for prepared executions (0: no limit) */
    uint32_t max_inflight;  /* concurrent queries on the device (1..16) */
    uint32_t reserved;
} kz_server_cfg;
#define KZ_QUERY_CLIQUE 1   /* k-clique count, k = 3..5 */
#define KZ_QUERY_CENSUS 2   /* motif census of a symmetric graph */
#define KZ_PENDING 1        /* kz_server_poll: the query has not finished */
typedef struct kz_query_result {
    uint64_t counts[KZ_CENSUS_PATTERNS]; /* clique: counts[0] */
    double queue_ms;  /* submit -> launch */
    double run_ms;    /* launch -> result on the host */
    double total_ms;  /* submit -> result on the host */
    int status;       /* KZ_OK or the query's error code */
    int reserved;
} kz_query_result;
int kz_server_start(const kz_server_cfg* cfg, kz_server* out);
int kz_server_submit(kz_server s, kz_graph g, int kind, int k, uint64_t* ticket);
/* KZ_OK: done, *out filled and the ticket retired; KZ_PENDING: not done
 * (wait == 0 only); wait != 0 blocks until the query finishes.  For a
 * query that failed (out->status != KZ_OK), kz_last_error() in the calling
 * thread then returns the failure's message. */
int kz_server_poll(kz_server s, uint64_t ticket, int wait, kz_query_result* out);
/* Finishes every submitted query, stops the worker, releases the prepared
 * executions the server holds and frees the server. */
int kz_server_stop(kz_server s);

/* ---- Set primitives, batched -------------------------------------------
 * Many (A, B) list pairs in one launch, one wave64 per pair.  Lists are
 * packed in `values` with offsets list_off[0..num_lists] (strictly
 * increasing ids per list); pair i uses lists a_idx[i] and b_idx[i].
 *   KZ_SETOP_INTERSECT   VertexSet::intersect_with (src/vertex_set.cc:20-49)
 *   KZ_SETOP_SUBTRACT    VertexSet::subtracted_by  (src/vertex_set.cc:51-86)
 *     result i is written sorted at out_values + out_off[i]; the slot must
 *     hold min(|A|,|B|) (intersect) or |A| (subtract) ids; out_len[i] gets
 *     the result length;
 *   KZ_SETOP_INTERSECT_COUNT  VertexSet::intersect_and_count
 *     (src/vertex_set.cc:88-111): out_len[i] = |A & B|;
 *   KZ_SETOP_TRIM_LOWER_BOUND trim_v_list_lower_bound (include/utilities.h:
 *     76-86): out_len[i] = index of the first id of A above bounds[i].
 * Host buffers in and out; b_idx is ignored by TRIM, bounds by the others. */
#define KZ_SETOP_INTERSECT 1
#define KZ_SETOP_SUBTRACT 2
#define KZ_SETOP_INTERSECT_COUNT 3
#define KZ_SETOP_TRIM_LOWER_BOUND 4
int kz_set_op_batch(int op, const uint32_t* values, const uint64_t* list_off, uint32_t num_lists,
                    const uint32_t* a_idx, const uint32_t* b_idx, const uint32_t* bounds, uint32_t num_pairs,
                    uint32_t* out_values, const uint64_t* out_off, uint32_t* out_len);

/* kz_set_op_batch without any device allocation: no reference counterpart.
 * For callers that cannot allocate device memory (tight serving loops,
 * harnesses that forbid hipMalloc).  Same arguments, checks and results;
 * the device staging of the batch (lists, offsets, pair indices, results:
 * the sum of each region rounded up to 256 bytes) lives in a file-scope
 * device array of KZ_SETOP_NOALLOC_MAX_BYTES that each device the library
 * runs on holds until the process ends.  A larger batch fails with
 * KZ_ERR_INVALID before any device work.  Calls share the array: the
 * library serializes them within a process; each call keeps its staging
 * until the next call overwrites it. */
#define KZ_SETOP_NOALLOC_MAX_BYTES (32ull << 20)
int kz_set_op_batch_noalloc(int op, const uint32_t* values, const uint64_t* list_off, uint32_t num_lists,
                            const uint32_t* a_idx, const uint32_t* b_idx, const uint32_t* bounds, uint32_t num_pairs,
                            uint32_t* out_values, const uint64_t* out_off, uint32_t* out_len);

/* ---- Plan IR: a generated apply() schedule in data form ----------------
 * Replaces the ApplyFunction callback (include/engine.h:589) with a table the
 * GPU executes.  An embedding of size s (matched vertices m_0..m_{s-1}) is
 * extended by the step list selected by its ORDER CODE: with r_j the dense
 * rank of m_j among the distinct matched values, code = sum_j r_j * s^j.
 * A step list runs up to KZ_MAX_OPS set operations (intersect_with /
 * subtracted_by, src/vertex_set.cc:20-86) whose result j is SLOT(s, j), then
 * up to KZ_MAX_OUTS outputs that iterate a set: SCATTER makes a child of
 * size s+1 per passing candidate (EmbeddingExplorationEngine::scatter,
 * src/engine.cc:1385-1453), COUNT adds the number of passing candidates to an
 * aggregator (Aggregator::add, include/engine.h:649-651).
 * A candidate v passes when bit p of `pred` is set, where its POSITION CLASS
 * p = 2 * |{distinct matched values < v}| + [v equals a matched value]. */
#define KZ_SET_NONE 0
#define KZ_SET_NBR 1  /* N(m_a)                                          */
#define KZ_SET_SLOT 2 /* slot b computed by the step list at size a <= s */
typedef struct kz_set_ref {
    uint8_t kind, a, b, reserved;
} kz_set_ref;

#define KZ_OP_INTERSECT 1
#define KZ_OP_SUBTRACT 2
typedef struct kz_op {
    uint8_t op, reserved[3];
    kz_set_ref lhs, rhs;
} kz_op;

#define KZ_OUT_SCATTER 1
#define KZ_OUT_COUNT 2
typedef struct kz_out {
    uint8_t kind;
    uint8_t aggregator;   /* COUNT target                                  */
    uint8_t child_active; /* SCATTER: is_new_v_active (informational)      */
    uint8_t finalized;    /* SCATTER: is_finalized (informational)         */
    kz_set_ref src;
    uint32_t pred;
} kz_out;

typedef struct kz_step_list {
    uint8_t num_ops, num_outs, reserved[2];
    kz_op ops[KZ_MAX_OPS];
    kz_out outs[KZ_MAX_OUTS];
} kz_step_list;

typedef struct kz_plan_desc {
    uint32_t pattern_size;    /* K: embeddings of size 1..K-1 are extended */
    uint32_t num_aggregators;
    uint32_t multiplicity[KZ_MAX_AGGREGATORS]; /* informational: printed count = raw / multiplicity */
    uint32_t num_step_lists;
    kz_step_list lists[KZ_MAX_STEP_LISTS];
    /* type_table[s][code] = index into lists, or 0xFF if that order type is
     * never reached.  Codes are < s^s <= 256. */
    uint8_t type_table[KZ_MAX_PATTERN][256];
} kz_plan_desc;

/* Built-in plans: the kautomine schedules of triangle.cc (k=3),
 * four_clique.cc (k=4), five_clique.cc (k=5). */
int kz_plan_clique(int k, kz_plan_desc* out);

/* Check a plan for structural validity (references to slots that exist and
 * are computed before use, sizes in range, last size has no scatter). */
int kz_plan_validate(const kz_plan_desc* plan);

/* Execute a plan with the plain level-by-level form.  raw_counts[a] gets the
 * uint64 sum of aggregator a (mod 2^64), before multiplicity division, as
 * Aggregator<EdgeId>::evaluate returns it (include/engine.h:628-645). */
int kz_count_plan(kz_graph g, const kz_plan_desc* plan, uint64_t* raw_counts, uint32_t num_counts,
                  kz_run_stats* stats);

/* Per-level HBM budget of the plain form, in bytes (the analogue of
 * memory_size_per_level, src/distributed_application.cc:20-22).  A smaller
 * budget forces more chunks; results do not change. */
int kz_set_level_budget(uint64_t bytes);

#ifdef __cplusplus
}
#endif

#endif /* KZ_H */
