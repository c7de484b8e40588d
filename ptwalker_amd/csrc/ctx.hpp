// The context behind the C ABI's opaque ptw_ctx, shared by api.cpp (graphs,
// single walks, buffers) and batch_api.cpp (batched queries).
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/ptw.h"
#include "engine.hpp"

namespace ptw {

// Device pools and tables of the batched-query path.  The slice pools are
// kept clean between batches (dense counters zero, hash slots empty), so a
// batch only has to clean what the previous one touched.
struct BatchState {
    uint32_t *dense = nullptr;
    uint64_t dense_cap = 0;  // u32 counters
    uint2 *hash = nullptr;
    uint64_t hash_cap = 0;  // slots
    uint2 *touched = nullptr;  // (device id, count) of each query's visited vertices
    uint64_t touched_cap = 0;
    BatchQuery *queries = nullptr;
    uint32_t *touched_count = nullptr, *overflow = nullptr;
    unsigned long long *steps = nullptr;
    double *entropy = nullptr;
    TopkStats *tk_stats = nullptr;
    uint32_t q_cap = 0;
    unsigned long long *tk_cand = nullptr;
    uint64_t cand_cap = 0;
    BatchChunk *chunks = nullptr;
    uint64_t chunk_cap = 0;
    uint32_t *topk_v = nullptr;
    unsigned long long *topk_c = nullptr;
    uint64_t topk_cap = 0;
    unsigned long long *hdr = nullptr;    // chunks, queries, k, seed, semantics, weighted, dense chunks, log pairs
    unsigned long long *queue = nullptr;  // chunk queue heads, log segment head, log tasks
    // visit log of logged dense queries (BatchLaunch::vlog and the arrays after it)
    uint32_t *vlog = nullptr, *seg_q = nullptr, *seg_n = nullptr, *seg_list = nullptr;
    uint16_t *part16 = nullptr;
    uint64_t log_segs = 0;  // segments allocated
    uint32_t *qlog = nullptr, *log_order = nullptr;
    unsigned long long *q_base = nullptr;
    uint32_t *part_cnt = nullptr, *part_loc = nullptr, *part_cur = nullptr, *run_loc = nullptr, *run_len = nullptr;
    uint64_t log_q_cap = 0;   // queries the per-query log arrays hold
    uint64_t log_rows = 0;    // (query, bucket) entries the per-bucket arrays hold
    uint32_t log_shift = 0, log_nb = 0;  // of the last plan with logged queries
    bool log_active = false;  // the current batch has logged queries (its launch runs the log's kernels)
    bool dirty = false;                   // the last batch's slices are not yet cleaned
    bool weighted = false;
    uint32_t last_nq = 0;
    std::vector<BatchQuery> h_queries;
    std::vector<BatchChunk> h_chunks;
    std::vector<uint32_t> h_log_order;  // logged queries, largest expected log first
    unsigned long long h_hdr[kBatchHdrWords] = {};
    // (query, vertex) gathers of the last batch, grown and kept
    uint32_t *g_q = nullptr, *g_v = nullptr;
    uint64_t *g_out = nullptr;
    uint64_t g_cap = 0;
    // captured launch sequence, valid while the pools and the graph keep their addresses
    hipGraphExec_t exec = nullptr;
};

// A batch launch pipeline of the serving loop: its own pools, stream and
// events, so two batches can be on the device at once, and page-locked
// staging that the batch's results are copied into without a wait.
struct ServeLane {
    BatchState b;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    hipEvent_t done = nullptr;  // after the result copies of the lane's batch
    unsigned long long *h_steps = nullptr;
    uint32_t *h_over = nullptr;
    double *h_ent = nullptr;
    uint32_t *h_tv = nullptr;
    unsigned long long *h_tc = nullptr;
    uint64_t h_q = 0, h_k = 0;  // staging capacity: queries, top-k entries
};
constexpr int kServeLanes = 2;

}  // namespace ptw

struct ptw_ctx {
    int device = 0;
    int num_cus = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    std::string err;

    bool has_graph = false;
    ptw::HostCsr host;
    ptw::DeviceGraph dev;
    uint32_t max_degree = 0, max_degree_vertex = 0, num_dangling = 0;

    // host copies of the relabel (empty without one)
    std::vector<uint32_t> h_perm, h_inv, h_drow;
    // weighted sampling: weights and alias tables in graph-id CSR order
    std::vector<float> weights;
    std::vector<uint32_t> alias_thr, alias_idx;

    unsigned long long *counts_dev = nullptr;  // V device-id counters
    uint64_t *counts_stage = nullptr;          // V graph-id counts staged for a host output
    uint64_t *pinned = nullptr;                // host, page-locked: [0, 8) scratch words, [8, 8 + V) counts
    unsigned long long *scratch = nullptr;     // 8 words
    // this device's address of the no-allocation walk's static scratch
    // (looked up at ptw_create with the device current)
    unsigned long long *static_scratch = nullptr;
    ptw::LaunchShape shape;
    ptw::VlogBuffers vlog;  // visit log of count walks, grown on demand
    ptw::BatchState batch;
    ptw::ServeLane serve[ptw::kServeLanes];  // created by the first ptw_serve_replay
};

namespace ptw {

inline int fail(ptw_ctx *ctx, int code, const std::string &msg)
{
    if (ctx) ctx->err = msg;
    return code;
}

#define HIP_TRY(ctx, call)                                                                        \
    do {                                                                                          \
        hipError_t e_ = (call);                                                                   \
        if (e_ != hipSuccess)                                                                     \
            return ::ptw::fail((ctx), e_ == hipErrorOutOfMemory ? PTW_ERR_OOM : PTW_ERR_HIP,      \
                               std::string(#call) + ": " + hipGetErrorString(e_));                \
    } while (0)

template <typename T>
int upload(ptw_ctx *ctx, T **dst, const T *src, size_t n)
{
    HIP_TRY(ctx, hipMalloc((void **)dst, (n ? n : 1) * sizeof(T)));
    if (n) HIP_TRY(ctx, hipMemcpy(*dst, src, n * sizeof(T), hipMemcpyHostToDevice));
    return PTW_OK;
}

// Builds the graph's neighbour records (weighted: the pairs for path walks)
// on the context's stream and waits for them; false if the memory is not
// there.  Kept with the graph once built.
bool ensure_edge_records(ptw_ctx *ctx, bool weighted);
// Frees the batch pools, the serving lanes' included (graph change or
// context teardown).
void free_batch(ptw_ctx *ctx);
// Destroys the serving lanes' streams, events and staging (context teardown).
void destroy_serve_lanes(ptw_ctx *ctx);

// Where a batch runs: its pools and tables, the stream its work is queued
// on, the events around its device work, and the CUs its persistent walk
// kernels fill (all of them, or a share when lanes run side by side; the
// other kernels are ordinary grids sized for the whole device).
struct LaneRef {
    BatchState &b;
    hipStream_t s;
    hipEvent_t ev0, ev1;
    int cus;
};
inline LaneRef main_lane(ptw_ctx *ctx) { return LaneRef{ctx->batch, ctx->stream, ctx->ev0, ctx->ev1, ctx->num_cus}; }

// ptw_batch_ppr's checks, plan, admission and device work on a lane, queued
// without a wait; the caller queues the result copies and waits.
// allow_log = false plans dense queries with slices even when visit_log is on.
int batch_enqueue(ptw_ctx *ctx, LaneRef L, const ptw_batch_desc &d, bool allow_log = true);
// Queues the copies of the lane's last batch results into host memory
// (page-locked for an asynchronous copy); any pointer may be null.
int batch_enqueue_results(ptw_ctx *ctx, LaneRef L, uint32_t nq, uint32_t k, unsigned long long *steps,
                          uint32_t *over, double *ent, uint32_t *tv, unsigned long long *tc);
// After the lane's batch completed: hash overflows and u32 count totals.
int batch_check(ptw_ctx *ctx, const ptw_batch_desc &d, const unsigned long long *steps, const uint32_t *over);
// Counts of (query, graph id) pairs of the lane's last batch (synchronous).
int batch_gather_lane(ptw_ctx *ctx, LaneRef L, uint64_t n, const uint32_t *query_index, const uint32_t *vertices,
                      uint64_t *out);

// Drops the captured batch graph (its kernels' arguments or launch shape changed).
inline void drop_batch_graph(ptw_ctx *ctx)
{
    if (ctx->batch.exec) {
        if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
        (void)hipGraphExecDestroy(ctx->batch.exec);
    }
    ctx->batch.exec = nullptr;
}

// Continuation threshold of a termination probability: continue iff r.z < thr.
inline uint32_t continue_threshold(double tp)
{
    double c = __builtin_floor((1.0 - tp) * 4294967296.0);
    if (c < 0.0) c = 0.0;
    if (c > 4294967295.0) c = 4294967295.0;
    return (uint32_t)c;
}

}  // namespace ptw
