/*
 * ptw.h -- C ABI of the MI355X probabilistic-termination random-walk engine.
 *
 * One context owns one HIP device, one stream and one resident graph.  Every
 * function returns 0 on success or a negative PTW_ERR_* code; the message of
 * the last failure is available from ptw_last_error().  No C++ exception
 * crosses this boundary.  A context is not thread-safe: use one per host
 * thread.
 *
 * Vertex ids at this boundary are "graph ids": for a loaded edge list they
 * are the first-appearance compact ids that the reference assigns while
 * loading (src/core/graph.hpp:293-319); for a CSR handed in they are the
 * CSR's own indices.  The device holds a degree-descending relabel of the
 * graph, which is invisible here: every output is mapped back.
 *
 * Walk process (one Philox4x32-10 block per step, keyed by seed and counted
 * by (walker id, step, stream)) -- see ptw_walk().
 */
#ifndef PTW_H
#define PTW_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PTW_VERSION_MAJOR 0
#define PTW_VERSION_MINOR 1

/* status codes */
#define PTW_OK 0
#define PTW_ERR_INVALID_ARG (-1)
#define PTW_ERR_IO (-2)
#define PTW_ERR_HIP (-3)
#define PTW_ERR_NO_GRAPH (-4)
#define PTW_ERR_OOM (-5)
#define PTW_ERR_DANGLING (-6)
#define PTW_ERR_FORMAT (-7)
#define PTW_ERR_UNSUPPORTED (-8)
#define PTW_ERR_ADMISSION (-9)
#define PTW_ERR_CANCELLED (-10)
#define PTW_ERR_TIMEOUT (-11)

/* graph formats: include/type.hpp:12-15 (BinaryGraphFormat, TextGraphFormat) */
#define PTW_FORMAT_BINARY 0
#define PTW_FORMAT_TEXT 1

/* apps: src/core/{pagerank,ppr,RWR,deepwalk}.cpp */
#define PTW_APP_PAGERANK 0
#define PTW_APP_PPR 1
#define PTW_APP_RWR 2
#define PTW_APP_DEEPWALK 3
/* Second-order (node2vec) walks of walk_len vertices: from cur, reached
 * from prev, a neighbour x is taken with weight 1/p if x == prev, 1 if x is
 * a neighbour of prev, 1/q otherwise; the first step is uniform.  The
 * reference declares the app (src/core/node2vec.cpp, flags -p -q in
 * include/option.hpp:447-482) but its sampler is not wired in
 * (src/core/solver.hpp:155-159).  Decisions: up to 64 proposals, attempt a
 * drawing Philox(seed; wid, step + a * 2^20, stream), slot j = mulhi64(r.x,
 * r.y, deg + k), accepted iff r.z < the slot's weight; then one exact draw
 * from the integer weights (64-bit r.y:r.x, attempt 64).  Weights are
 * floor(2^32 w(x) / M), M = max(1/p, 1, 1/q), and k = 0.  On a graph where
 * no row repeats a neighbour and every edge has its reverse, with 1/p above
 * max(1, 1/q): M = max(1, 1/q), prev's own slot accepts always, and its
 * excess e = floor(2^32 / (p M)) - 2^32 rides on k = ceil(e / 2^32) extra
 * slots accepting prev below floor(e / k).  With p = q = 1 the paths are
 * DEEPWALK's.  The first node2vec walk on a graph builds an edge set of
 * 8 B x (the smallest power of two >= 2E) on the device. */
#define PTW_APP_NODE2VEC 4

/* step-budget semantics for the probabilistic-termination apps */
#define PTW_SEM_REFERENCE 0 /* L = 1 + Geometric(tp), src/core/pagerank.cpp:44-67 */
#define PTW_SEM_TEXTBOOK 1  /* L = Geometric(tp) (may be 0) */

/* start vertex of a walker */
#define PTW_START_UNIFORM 0    /* uniform over vertices, src/core/pagerank.cpp:70-80 */
#define PTW_START_SOURCE 1     /* every walker at desc.source, src/core/ppr.cpp:91-108 */
#define PTW_START_ROUNDROBIN 2 /* walker w at vertex w mod V ("walks per vertex") */

/* what a walker does at a vertex without out-edges while budget is left */
#define PTW_DANGLING_ERROR 0 /* report PTW_ERR_DANGLING (the reference divides by 0) */
#define PTW_DANGLING_STOP 1  /* the walker ends there */

/* ptw_walk_desc.flags */
#define PTW_FLAG_OUT_DEVICE 1u /* output pointers are device pointers */
#define PTW_FLAG_ACCUMULATE 2u /* counts += instead of counts = */
#define PTW_FLAG_WEIGHTED 4u   /* neighbour choice by edge weight (alias tables) */

/* ptw_graph_rmat flags */
#define PTW_RMAT_KEEP_ISOLATED 1u /* keep degree-0 ids (default: compact them away) */
#define PTW_GRAPH_NO_RELABEL 2u   /* keep graph ids on the device (no degree relabel) */
#define PTW_GRAPH_INGEST_HOST 4u  /* ptw_graph_load: parse and build the CSR on the host */
#define PTW_GRAPH_INGEST_DEVICE 8u /* ptw_graph_load: parse and build the CSR on the device */

#define PTW_INVALID_VERTEX 0xFFFFFFFFu

/* ptw_create device argument for a context without a device: graph ingest,
 * export and writing work; walks return PTW_ERR_UNSUPPORTED. */
#define PTW_DEVICE_HOST_ONLY (-1)

typedef struct ptw_ctx ptw_ctx;

typedef struct ptw_graph_info {
    uint32_t num_vertices;
    uint32_t max_degree;
    uint64_t num_edges;
    uint32_t max_degree_vertex; /* graph id; lowest id among ties */
    uint32_t num_dangling;      /* vertices with out-degree 0 */
    uint32_t hub_vertices;      /* configured LDS hub counters (min(V, lds_hubs)); the count a
                                   walk used is in ptw_walk_out.launch_lds_hubs */
    uint32_t relabelled;        /* 1 if the device holds a degree relabel */
    uint64_t device_bytes;      /* HBM held by the resident graph */
    uint32_t weighted;          /* 1 if edge weights and alias tables are resident */
    uint32_t reserved;
} ptw_graph_info;

/* Plain-data replacement for the reference's Alg template concept
 * (src/core/solver.hpp:1049-1050; type/walking_step/walker_init/update,
 * src/core/pagerank.cpp:10-90). */
typedef struct ptw_walk_desc {
    int32_t app;             /* PTW_APP_* */
    int32_t semantics;       /* PTW_SEM_* */
    int32_t start_mode;      /* PTW_START_* */
    int32_t dangling_policy; /* PTW_DANGLING_* */
    uint32_t source;         /* graph id, PTW_START_SOURCE */
    uint32_t walk_len;       /* DEEPWALK, NODE2VEC: vertices per path (reference -l) */
    double tp;               /* termination probability (reference --sp), (0, 1] */
    uint64_t seed;           /* Philox key */
    uint64_t wid_base;       /* first walker id (for split or resumed runs) */
    uint64_t walkers;        /* walker count (reference -w) */
    uint32_t stream;         /* 4th Philox counter word (query id) */
    uint32_t flags;          /* PTW_FLAG_* */
    double p;                /* NODE2VEC return parameter, [2^-16, 2^16] (reference -p) */
    double q;                /* NODE2VEC in-out parameter, [2^-16, 2^16] (reference -q) */
} ptw_walk_desc;

typedef struct ptw_walk_out {
    uint64_t *counts; /* PAGERANK/PPR: u64[V] visit counts, or NULL */
    uint32_t *ends;   /* RWR: u32[walkers] end vertex per walker, or NULL */
    uint32_t *paths;  /* DEEPWALK, NODE2VEC: u32[walkers * walk_len], or NULL */
    uint32_t *starts; /* RWR, optional: u32[walkers] start vertex per walker, so (starts[w],
                         ends[w]) is the reference's {start_v, end_v} pair (src/core/RWR.cpp:9-13,
                         76-89, 99-105); NULL for the other apps */
    /* filled on return */
    uint64_t steps;    /* sum of transitions over walkers */
    uint64_t visits;   /* count increments (count apps) */
    uint64_t dangling; /* walkers stopped at a degree-0 vertex */
    double kernel_ms;  /* device time of the walk kernel and, for count apps, of the
                          visit log's counting passes (HIP events) */
    double total_ms;   /* wall time of the call */
    /* the launch the walk ran with, after the automatic rules of
     * ptw_tuning (blocks_per_cu = 0, cache_slots = PTW_CACHE_AUTO) */
    uint32_t launch_block;       /* threads per workgroup */
    uint32_t launch_blocks;      /* workgroups */
    uint32_t launch_chunk;       /* walker ids per queue take */
    uint32_t launch_lds_hubs;    /* u16 LDS hub counters (count apps) */
    uint32_t launch_cache_slots; /* LDS visit-cache slots (count apps) */
    uint32_t launch_visit_log;   /* visit-log entries written (count apps; 0: the walk
                                    counted with device atomics only) */
} ptw_walk_out;

/* ---- context ---------------------------------------------------------- */
/* Replaces FMobSolver(Graph*, MultiThreadConfig), src/core/solver.hpp:120,
 * and init_concurrency, include/numa_helper.hpp:37-127. */
int ptw_create(ptw_ctx **out, int device); /* device >= 0, or PTW_DEVICE_HOST_ONLY */
void ptw_destroy(ptw_ctx *ctx);
const char *ptw_last_error(const ptw_ctx *ctx);
/* Library version as major * 1000 + minor. */
int ptw_version(void);

/* ---- graph ------------------------------------------------------------ */
/* Graph::load(path, GraphFormat, as_undirected, bodary),
 * src/core/graph.hpp:277-366, with read_text_graph / read_binary_graph,
 * include/io.hpp:39-53,76-90.  Ids are compacted by first appearance and
 * adjacency keeps file order.  flags: PTW_GRAPH_NO_RELABEL, and where the
 * file is parsed and the CSR built: PTW_GRAPH_INGEST_HOST (threads over an
 * mmap) or PTW_GRAPH_INGEST_DEVICE (the file is copied to HBM, parsed and
 * compacted there, and the CSR comes back); by default the device for files
 * of 16 MiB and more on a device context.  Both give the same graph. */
int ptw_graph_load(ptw_ctx *ctx, const char *path, int format, int as_undirected, uint32_t flags);

/* A CSR in graph ids (row_ptr[V + 1], col[E]); replaces the graph build of
 * src/core/graph.hpp:331-366 for callers that already hold a CSR. */
int ptw_graph_from_csr(ptw_ctx *ctx, const uint32_t *row_ptr, const uint32_t *col,
                       uint32_t num_vertices, uint64_t num_edges, uint32_t flags);

/* Seeded R-MAT generator: 2^scale vertices, edge_factor * 2^scale draws with
 * quadrant probabilities a, b, c (d = 1 - a - b - c); symmetrized,
 * deduplicated, self-loops dropped, isolated ids compacted away unless
 * PTW_RMAT_KEEP_ISOLATED.  Replaces gen_undirected_graph of
 * src/test/test.hpp:62-88 as the synthetic-input source. */
int ptw_graph_rmat(ptw_ctx *ctx, uint32_t scale, uint32_t edge_factor, double a, double b,
                   double c, uint64_t seed, uint32_t flags);

/* Edge weights for weighted neighbour choice, one per CSR position in graph
 * ids (finite, > 0).  The reference samples uniformly only (AdjUnit holds no
 * weight, src/core/graph.hpp:29-31); weighted walks are an addition.
 * Weights are quantized per vertex to integers q = floor(w / w_max 2^24 +
 * 0.5) clamped to [1, 2^24], and each vertex gets an alias table built by
 * Vose's method in exact integer arithmetic: a step picks slot
 * j = mulhi64(r.xy, deg) and takes neighbour j if r.w < thr[j], else
 * alias[j]. */
int ptw_graph_set_weights(ptw_ctx *ctx, const float *weights, uint64_t num_edges);
/* U(0, 1] weights on a 2^-24 grid from Philox(seed, edge index). */
int ptw_graph_random_weights(ptw_ctx *ctx, uint64_t seed);
/* Copy weights and alias tables (graph-id CSR order); any pointer may be NULL. */
int ptw_graph_export_weights(ptw_ctx *ctx, float *weights, uint32_t *alias_thr, uint32_t *alias_idx);

int ptw_graph_info_get(ptw_ctx *ctx, ptw_graph_info *info);

/* The id each graph id had before compaction, names[V]: the file's own id
 * for a loaded edge list (the inverse of the first-appearance map, which the
 * reference builds in src/core/graph.hpp:293-319 and leaves unexported,
 * id2name "TBD" at :663-665), the generator's id for R-MAT, the CSR index for
 * a CSR handed in.  Maps every output (counts, ends, starts, paths, top-k)
 * back to the input's ids. */
int ptw_graph_export_names(ptw_ctx *ctx, uint32_t *names);

/* Copy the resident graph's CSR in graph ids to host memory. */
int ptw_graph_export_csr(ptw_ctx *ctx, uint32_t *row_ptr, uint32_t *col);

/* Write the resident graph as an edge list (write_text_graph /
 * write_binary_graph, include/io.hpp:55-105). */
int ptw_graph_write(ptw_ctx *ctx, const char *path, int format);

/* ---- walks ------------------------------------------------------------ */
/* walk(FMobSolver*, walker_num, walk_len, mem_quota, Alg&, walker_size),
 * src/core/solver.hpp:1049-1074, with the walker loop of
 * src/core/solver.hpp:194-498.  Process for walker w (wid = wid_base + w):
 *   start:  SOURCE: desc.source; ROUNDROBIN: wid mod V; UNIFORM:
 *           mulhi64(Philox(wid, 0xFFFFFFFF, stream).xy, V)
 *   step i: r = Philox(ctr = (wid_lo, wid_hi, i, stream), key = seed);
 *           PT apps stop when r.z >= floor((1 - tp) 2^32) (from i >= 1 under
 *           PTW_SEM_REFERENCE); DEEPWALK stops at i == walk_len - 1;
 *           v = col[row_ptr[v] + mulhi64(r.xy, deg v)]
 * PAGERANK/PPR count the start and every step (type 0); RWR records the
 * last vertex (type 1); DEEPWALK writes the path (PTW_INVALID_VERTEX after a
 * dangling stop).  Synchronous: returns when outputs are complete. */
int ptw_walk(ptw_ctx *ctx, const ptw_walk_desc *desc, ptw_walk_out *out);

/* ptw_walk without any allocation, for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  No counterpart in the reference.  Same process and
 * results as ptw_walk; the outputs must be device buffers (PTW_FLAG_OUT_DEVICE).
 * The walk's scratch -- the queue and total words and, for PAGERANK/PPR, the
 * V device-id visit counters -- lives in a 32 MiB array in the library's
 * device code, which each GPU holds until the process ends: count walks take
 * graphs of at most 4,194,296 vertices (larger ones fail with
 * PTW_ERR_INVALID_ARG); NODE2VEC fails with PTW_ERR_UNSUPPORTED (its edge set
 * is built on first use).  The call allocates nothing, host memory included
 * (the totals come back through page-locked words made at graph install).
 * Every call on one GPU shares that array: calls that overlap on one GPU,
 * from any context or thread, race. */
int ptw_walk_static(ptw_ctx *ctx, const ptw_walk_desc *desc, ptw_walk_out *out);

/* ---- batched queries -------------------------------------------------- */
/* Many single-source queries in one persistent launch: PPR queries count
 * every visit (type 0, the start included), RWR queries count each walker's
 * last vertex (type 1, src/core/RWR.cpp:99-105), so their counts are the
 * endpoint histogram.  The reference runs one query per process
 * (src/core/ppr.cpp:121-139, src/core/RWR.cpp:109-128, one start vertex,
 * src/core/ppr.cpp:54-57); batching is an addition.  Query i's walker w
 * draws from Philox(seed; w, step, queries[i].stream), so its counts equal
 * those of ptw_walk(PPR or RWR, source, walkers, tp, stream) alone, in any
 * mix.  Per-query counters live in u32 slices: dense (V counters) for
 * walkers > dense_threshold, a hash table otherwise.  Walkers that reach a
 * vertex without out-edges stop there. */
#define PTW_QUERY_VISITS 0    /* PPR: visit counts */
#define PTW_QUERY_ENDPOINTS 1 /* RWR: end-vertex counts */

typedef struct ptw_query {
    uint32_t source;  /* graph id */
    uint32_t stream;  /* 4th Philox counter word of this query */
    uint64_t walkers; /* at most 2^32 / (1.5 c), c = 1 + 1/tp (visits) or 1 (endpoints); a
                         query whose count total reaches 2^32 (u32 counters may have
                         wrapped) fails with PTW_ERR_OOM */
    double tp;
    uint32_t kind;    /* PTW_QUERY_* */
    uint32_t reserved;
} ptw_query;

#define PTW_BATCH_GRAPH 1u    /* run the device sequence as a captured hipGraph (replayed across batches) */
#define PTW_BATCH_WEIGHTED 2u /* neighbour choice by edge weight */

typedef struct ptw_batch_desc {
    const ptw_query *queries;
    uint32_t num_queries;
    uint32_t topk;            /* results per query, <= 1024 (0: none) */
    uint64_t seed;
    int32_t semantics;        /* PTW_SEM_* */
    uint32_t flags;           /* PTW_BATCH_* */
    uint64_t dense_threshold; /* walkers above this are dense (workgroup chunks, the visit log or
                                 a dense slice); 0 = 1 (every query dense) when dense
                                 queries are logged (visit_log != 0), 100000 with dense
                                 slices */
} ptw_batch_desc;

typedef struct ptw_batch_out {
    uint32_t *topk_vertex; /* host [num_queries * topk]: graph ids, count desc (ties: degree desc, id asc) */
    uint64_t *topk_count;  /* host [num_queries * topk] visit counts (PTW_INVALID_VERTEX / 0 padding) */
    uint64_t *steps;       /* host [num_queries] or NULL */
    double *entropy;       /* host [num_queries] or NULL: entropy (nats) of each query's visit distribution */
    /* filled on return */
    uint64_t total_steps;
    uint64_t slice_bytes; /* HBM held by the counter slices of this batch */
    uint32_t dense_queries;
    uint32_t hash_queries;
    double kernel_ms; /* device time: counter setup, walk, top-k */
    double total_ms;
    /* dense queries counted through the visit log instead of a slice (visit_log
       tuning != 0); a batch whose log fills is run again with slices, which
       gives the same counts (log_rerun = 1) */
    uint32_t logged_queries;
    uint32_t log_rerun;
} ptw_batch_out;

int ptw_batch_ppr(ptw_ctx *ctx, const ptw_batch_desc *desc, ptw_batch_out *out);
/* HBM the batch would need beyond what the context already holds (admission). */
int ptw_batch_plan_bytes(ptw_ctx *ctx, const ptw_batch_desc *desc, uint64_t *bytes);
/* Full count vector (graph ids, host u64[V]) of query `index` of the last
 * batch: visits, or end vertices for an endpoint query. */
int ptw_batch_counts(ptw_ctx *ctx, uint32_t index, uint64_t *counts);
/* Writes the top-k lists of a finished batch as text, one line per entry:
 * "query source kind rank vertex count" (kind 0 visits, 1 endpoints; rank
 * from 0; padding entries are left out), after a "#" header line.  Vertex
 * ids are graph ids (PTW_IDS_GRAPH) or the input's own ids
 * (PTW_IDS_ORIGINAL, see ptw_graph_export_names).  desc and out are the
 * arguments of that batch's ptw_batch_ppr.  No counterpart in the reference,
 * which prints one count per process (src/core/pagerank.cpp:110). */
#define PTW_IDS_GRAPH 0
#define PTW_IDS_ORIGINAL 1
int ptw_batch_write_topk(ptw_ctx *ctx, const char *path, const ptw_batch_desc *desc, const ptw_batch_out *out,
                         int id_space);

/* Counts of n (query index, graph id) pairs of the last batch, on the device
 * in one pass: out[i] = count of vertices[i] in query query_index[i]. */
int ptw_batch_gather(ptw_ctx *ctx, uint64_t n, const uint32_t *query_index, const uint32_t *vertices,
                     uint64_t *out);

/* ---- FP64 solver ------------------------------------------------------ */
/* Expected visits per walker of single-source walks from each of n sources,
 * by FP64 power iteration on the device (pull over the transposed CSR, up to
 * 64 sources per pass):  x = e_s + sum_{i>=1} P(L >= i) e_s P^i  with
 * P(L >= i) = (1-tp)^(i-1) (PTW_SEM_REFERENCE) or (1-tp)^i (TEXTBOOK),
 * iterated until the remaining mass bound (1-tp)^i / tp is below tol.
 * Mass that reaches a vertex without out-edges stops there, as walkers do.
 * out: host double[n * V] in graph ids, or NULL.  topk_*: host [n * k] of
 * the k largest entries per source (value desc, id asc), or k = 0.
 * This is the FP64 oracle the walk estimates are scored against on the
 * device; no counterpart in the reference. */
int ptw_ppr_exact(ptw_ctx *ctx, const uint32_t *sources, const double *tps, uint32_t n, int32_t semantics,
                  double tol, double *out, uint32_t k, uint32_t *topk_vertex, double *topk_value,
                  uint32_t *iterations);

/* ---- budget predictor ------------------------------------------------- */
/* A 6-64-64-1 ReLU MLP that predicts log(walker budget) of a PPR query from
 * six features (log degree, log mean neighbour degree, log 2-hop degree sum,
 * pilot top-10 mass, pilot entropy, log target eps).  BF16 weights (round to
 * nearest even) with FP32 accumulation; forward, backward and Adam run as
 * HIP kernels; gradients are summed in a fixed order (deterministic).
 * Inputs and targets are standardized with the model's normalization.
 * No counterpart in the reference (its budget is the -w flag). */
typedef struct ptw_mlp ptw_mlp;
int ptw_mlp_create(ptw_ctx *ctx, uint64_t seed, ptw_mlp **out);
void ptw_mlp_destroy(ptw_mlp *mlp);
/* Number of FP32 parameters: W1[64][6], b1[64], W2[64][64], b2[64], W3[64], b3. */
int ptw_mlp_num_params(void);
int ptw_mlp_set_normalization(ptw_mlp *mlp, const float *mean6, const float *std6, float y_mean, float y_std);
/* params: FP32 master copy (host, ptw_mlp_num_params floats); norm: 14 floats
 * (mean6, std6, y_mean, y_std); either may be NULL. */
int ptw_mlp_get_params(ptw_mlp *mlp, float *params, float *norm);
int ptw_mlp_set_params(ptw_mlp *mlp, const float *params);
/* y[i] = predicted log budget of x[i * 6 .. i * 6 + 5] (host arrays). */
int ptw_mlp_forward(ptw_mlp *mlp, const float *x, uint32_t n, float *y);
/* Gradient of the mean squared error (standardized target) over all n rows. */
int ptw_mlp_gradients(ptw_mlp *mlp, const float *x, const float *t, uint32_t n, float *grad, float *loss);
/* Adam (beta 0.9 / 0.999) over shuffled mini-batches of `batch` rows. */
int ptw_mlp_train(ptw_mlp *mlp, const float *x, const float *t, uint32_t n, uint32_t epochs, uint32_t batch,
                  float lr, uint64_t seed, float *final_loss);

/* ---- serving ---------------------------------------------------------- */
/* A request loop over the resident graph: requests (source, target eps)
 * arrive at given times and are served by continuous batching -- every launch
 * carries the pilot walks of the requests that arrived since the previous
 * launch and the main walks of requests whose budgets are known, admitted
 * while their counter slices fit the HBM allowance; the rest wait for the
 * next launch.  A request with id r walks Philox streams r (main),
 * r + 2^30 (pilot) and r + 2^31 (audit), so its answer does not depend on
 * what it was batched with.  No counterpart in the reference (one query per
 * process). */
#define PTW_BUDGET_FIXED 0   /* budget = fixed_c2 / eps^2 for every request */
#define PTW_BUDGET_LEARNED 1 /* budget = margin * exp(mlp(features)); fixed on out-of-range features */

typedef struct ptw_request {
    uint32_t source;  /* graph id */
    uint32_t id;      /* < 2^30; selects the request's Philox streams and audit draw */
    float eps;        /* target top-k L1 error */
    uint32_t reserved;
    double arrival_s; /* seconds after the call starts */
} ptw_request;

typedef struct ptw_serve_config {
    double tp;
    uint64_t seed;
    int32_t budget_mode;   /* PTW_BUDGET_* */
    uint32_t topk;         /* results per request, <= 1024 */
    double fixed_c2;       /* fixed budget numerator (also the fallback) */
    double margin;         /* learned budget safety factor */
    uint64_t w_min, w_max; /* budget clip */
    float feat_lo[6];      /* learned mode: a feature outside [lo, hi] selects the fixed budget */
    float feat_hi[6];
    double audit_rate;     /* fraction of requests re-walked at 10x budget; a short one is replaced by it */
    uint32_t max_batch;    /* queries per launch (pilots + mains + audits) */
    uint32_t reserved;
    uint64_t hbm_limit;    /* bytes of counter slices a launch may hold; 0 = the device's free HBM */
} ptw_serve_config;

typedef struct ptw_request_result {
    uint64_t budget;   /* walkers of the main walk */
    double admit_s;    /* launch start of the main walk (s since the loop started) */
    double done_s;     /* completion (s since the loop started) */
    uint32_t fallback; /* 1 if the fixed budget was used in learned mode */
    uint32_t audited;
    uint32_t topped_up;
    uint32_t launches; /* launches the request took part in */
} ptw_request_result;

typedef struct ptw_serve_stats {
    double makespan_s;
    double queries_per_s;
    double p50_ms, p99_ms, mean_ms; /* latency: done - arrival */
    uint64_t walk_steps;
    uint32_t launches, fallbacks, audited, topped_up;
    uint32_t admission_waits; /* launches that left a ready request waiting for HBM */
    uint32_t reserved;
} ptw_serve_stats;

/* Serves n requests in real time (arrival_s relative to the call); mlp may be
 * NULL in fixed mode.  topk_vertex / topk_count: host [n * topk].
 * Two lanes, each with its own stream (default priority), counter pools and
 * half of the CUs for its persistent kernels, carry one batch each; a lane
 * that is idle takes the next batch at once, so a request that arrives while
 * a long batch runs is walked beside it.  The lanes overlap only on separate
 * hardware queues: HIP maps streams onto GPU_MAX_HW_QUEUES queues (default 4),
 * so a process holding more than one context should raise it (each context
 * holds up to three streams) before HIP starts. */
int ptw_serve_replay(ptw_ctx *ctx, ptw_mlp *mlp, const ptw_serve_config *cfg, const ptw_request *reqs, uint32_t n,
                     ptw_request_result *results, uint32_t *topk_vertex, uint64_t *topk_count,
                     ptw_serve_stats *stats);

/* ---- launch tuning ---------------------------------------------------- */
/* Launch shape of the walk kernels.  The reference's counterparts are its
 * walker container size --ws and hub degree boundary --db
 * (include/option.hpp:243-300) and the sampler cost catalogue of
 * src/core/mini_bmk.hpp:29-118; here the knobs are GPU launch parameters,
 * chosen by the autotuner (ptwalker_amd.tune). */
/* cache_slots value that lets the library choose per launch: a count walk
 * from one source of out-degree <= 2048 keeps a 512-slot visit cache; any
 * other count walk gives the cache's 4 KiB to 2048 more hub counters (up to
 * 81408); batched launches keep 512 slots.  ptw_walk_out reports the choice. */
#define PTW_CACHE_AUTO (-1)

typedef struct ptw_tuning {
    int32_t block;         /* threads per workgroup: multiple of 64, 64..1024 */
    int32_t blocks_per_cu; /* resident workgroups per CU; 0 = occupancy query, RWR walks then
                              use workgroups of at most 512 threads, and a walk under 8
                              walkers per resident lane runs one workgroup of at most 512
                              threads per CU with at most 8192 hub counters and at most 1024
                              cache slots (none for uniform starts) */
    int32_t chunk;         /* walker ids a wavefront takes from the queue at once (>= 64) */
    int32_t lds_hubs;      /* u16 LDS counters for the top-degree vertices, 0..81408 */
    int32_t cache_slots;   /* LDS visit-cache slots (power of two, <= 16384; 0 = off;
                              PTW_CACHE_AUTO, the default);
                              2 B * lds_hubs + 8 B * cache_slots <= 159 KiB (auto: 512 slots) */
    int32_t batch_chunk;   /* walker ids per queue entry in batched launches (>= 64) */
    int32_t topk_tile;     /* LDS candidate slots of the top-k kernel (power of two, 256..4096) */
    int32_t dense_chunk;   /* walker ids per workgroup chunk of a dense batched query, a multiple
                              of 1024 up to 2^20; 0 = from the batch's dense walkers */
    int32_t visit_log;     /* count apps: 1 (default) = visits that no LDS counter takes are
                              appended to a device log and counted by id bucket after the
                              walk (a u32 and a u16 buffer of 1.1x the expected visits, kept
                              by the context, at most 2^30 entries each; walks under 8 walkers per
                              resident lane skip it); N > 1, a multiple of 1024 = the same
                              with at most N entries; 0 = one device atomic per such visit.
                              Visits beyond a full log take device atomics.  Batches: with
                              visit_log != 0 (and at most 4096 id buckets of 2^13..2^15 ids),
                              dense queries keep a hub slice and log the rest of their visits
                              (u32 entry, u16 partition entry and one pair place per expected
                              visit, 1.25x, held by the batch pools); N > 1 caps that log at
                              N entries, and a batch whose log fills runs again with dense
                              slices.  Counts are the same either way */
    int32_t edge_records;  /* neighbour records, built with the graph (weighted ones and any
                              missing ones by the first walk that reads them) and kept with
                              it: per CSR slot the neighbour's id, row start,
                              degree and graph id (16 B per edge; weighted path walks: both
                              alias candidates, 32 B per edge), so a step is one dependent
                              read.  1 (default) = every walk (weighted count and endpoint
                              walks keep the weighted-edge records; a node2vec proposal
                              reads its candidate's record); -1 = DEEPWALK and NODE2VEC
                              walks only; 0 = never.  Results are the same */
} ptw_tuning;
int ptw_tuning_get(ptw_ctx *ctx, ptw_tuning *t);
int ptw_tuning_set(ptw_ctx *ctx, const ptw_tuning *t);

/* ---- device buffers --------------------------------------------------- */
/* Device memory on the context's device for callers without a HIP
 * allocator of their own (outputs with PTW_FLAG_OUT_DEVICE).  No counterpart
 * in the reference, whose outputs are host arrays owned by the app
 * (src/core/pagerank.cpp:23, src/core/RWR.cpp:29). */
int ptw_buffer_alloc(ptw_ctx *ctx, uint64_t bytes, void **dptr);
int ptw_buffer_free(ptw_ctx *ctx, void *dptr);
int ptw_buffer_upload(ptw_ctx *ctx, void *dptr, const void *host, uint64_t bytes);
int ptw_buffer_download(ptw_ctx *ctx, const void *dptr, void *host, uint64_t bytes);
/* Zero `bytes` (a multiple of 4) with a fill kernel on the context's stream. */
int ptw_buffer_zero(ptw_ctx *ctx, void *dptr, uint64_t bytes);
/* Free and total HBM of the context's device. */
int ptw_device_memory(ptw_ctx *ctx, uint64_t *free_bytes, uint64_t *total_bytes);

/* ---- measurement ------------------------------------------------------ */
/* Random-gather microbenchmark: `gathers` dependent-free random u32 loads
 * from a table of `table_bytes`, Philox indices.  Reports gathers/s and the
 * kernel time.  No counterpart in the reference: it is the roofline the walk
 * kernel is measured against. */
int ptw_bench_gather(ptw_ctx *ctx, uint64_t table_bytes, uint64_t gathers, uint64_t seed,
                     double *gathers_per_s, double *kernel_ms);
/* Walk-step yardstick on the resident graph: `steps` vertices drawn in
 * proportion to out-degree (the source of a uniform random edge: on a
 * symmetric graph, the stationary state of a walk), then one step from each
 * -- row_ptr[v], row_ptr[v + 1] and one uniformly picked col entry -- with
 * the steps independent of each other and their vertices read in order.
 * Reports steps/s and the kernel time: the memory system's rate for a step's
 * own accesses, without the dependence between a walker's steps.  No
 * counterpart in the reference. */
int ptw_bench_step(ptw_ctx *ctx, uint64_t steps, uint64_t seed, double *steps_per_s, double *kernel_ms);
/* Self-check of the device primitives behind graph ingest: exclusive scans
 * (u8 -> u64, u32, u64), the stable radix sort of (u32 key, u32 value) pairs
 * over the low `bits` (1..32) key bits, the flagged select and the u32 max,
 * each run on n pseudo-random items drawn from `seed` (keys in [0, 2^bits),
 * with runs of repeated keys) and compared element by element with a host
 * computation (std::stable_sort for the sort).  Returns PTW_OK, or
 * PTW_ERR_FORMAT with the first difference in ptw_last_error.  n < 2^32 - 1.
 * No counterpart in the reference (which sorts on the host). */
int ptw_check_primitives(ptw_ctx *ctx, uint64_t n, int bits, uint64_t seed);
/* Streaming copy of `bytes` (multiple of 16) between two device buffers:
 * the stream roofline, and the known-size kernel that calibrates HBM byte
 * counters.  No counterpart in the reference. */
int ptw_bench_copy(ptw_ctx *ctx, uint64_t bytes, double *gb_per_s, double *kernel_ms);

#ifdef __cplusplus
}
#endif

#endif /* PTW_H */
