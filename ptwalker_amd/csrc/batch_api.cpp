// C ABI of batched queries: planning, pools, the launch sequence (eager or
// as one captured hipGraph replayed across batches), top-k and export.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>

#include "../../include/ptw.h"
#include "ctx.hpp"
#include "engine.hpp"

namespace ptw {

namespace {

double now_ms()
{
    using namespace std::chrono;
    return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

template <typename T>
void hfree(T *&p)
{
    if (p) (void)hipFree(p);
    p = nullptr;
}

void drop_exec(BatchState &b)
{
    if (b.exec) (void)hipGraphExecDestroy(b.exec);
    b.exec = nullptr;
}

void free_log(BatchState &b)
{
    hfree(b.vlog);
    hfree(b.seg_q);
    hfree(b.seg_n);
    hfree(b.seg_list);
    hfree(b.part16);
    hfree(b.qlog);
    hfree(b.log_order);
    hfree(b.q_base);
    hfree(b.part_cnt);
    hfree(b.part_loc);
    hfree(b.part_cur);
    hfree(b.run_loc);
    hfree(b.run_len);
    b.log_segs = b.log_q_cap = b.log_rows = 0;
    b.log_shift = b.log_nb = 0;
    b.log_active = false;
}

struct Plan {
    std::vector<BatchQuery> q;
    std::vector<BatchChunk> chunks;  // workgroup chunks of dense queries, then wavefront chunks of hash queries
    uint64_t dense = 0, hash = 0, touched = 0, cand = 0;  // elements needed
    uint32_t n_dense = 0, n_hash = 0;
    uint64_t dense_chunks = 0;
    // dense queries are logged (SLICE_LOG): buckets of 2^log_shift ids, log_nb of them
    bool log = false;
    uint32_t log_shift = 0, log_nb = 0;
    uint64_t log_entries = 0;  // expected log entries, 1.25x (the pair lists' rule)
    uint64_t log_hubs = 0;     // hub counters + 1 per logged query
    uint64_t log_segs = 0;     // segments the log may claim
    std::vector<uint32_t> log_order;  // logged queries, largest expected log first
    uint64_t touched_total() const { return touched + (log ? log_segs * kBatchLogSeg + log_hubs : 0); }
};

// Walker ids per workgroup chunk of a dense query: about 1/48 of the
// query's walkers, so that the resident workgroups (one per CU) share a
// handful of queries at any time and those queries' slices stay in the
// Infinity Cache; at least 16384 (16 walkers per lane of a 1024-thread
// workgroup, so the drain at the end of a chunk stays short) and at most
// 49152 (a large query gains nothing from longer chunks).
uint64_t dense_chunk_size(uint64_t walkers)
{
    const uint64_t c = (walkers / 48 + 4095) / 4096 * 4096;
    return std::min<uint64_t>(std::max<uint64_t>(c, 16384), 49152);
}

int make_plan(ptw_ctx *ctx, const ptw_batch_desc &d, Plan &P, bool allow_log = true)
{
    const uint32_t V = ctx->dev.V;
    const uint64_t chunk = (uint64_t)ctx->shape.batch_chunk;
    // dense queries are logged unless visit_log is 0 or the graph needs more
    // id buckets than the log's kernels hold
    const bool log_ok = allow_log && ctx->shape.visit_log && batch_log_fits(V);
    const uint64_t thr_dense =
        d.dense_threshold ? d.dense_threshold : log_ok ? kDenseThresholdLogged : kDenseThresholdSlices;
    P.log_shift = log_ok ? kBatchLogShift : 0u;
    P.log_nb = log_ok ? ((V - 1) >> kBatchLogShift) + 1 : 0u;
    const uint32_t log_hub_len =
        log_ok ? std::min<uint32_t>(batch_log_hubs(ctx->shape.lds_hubs, P.log_nb, batch_cache_slots(ctx->shape)),
                                    dense_pitch(V))
               : 0u;
    P.q.resize(d.num_queries);
    for (uint32_t i = 0; i < d.num_queries; ++i) {
        const ptw_query &qq = d.queries[i];
        if (qq.source >= V) return fail(ctx, PTW_ERR_INVALID_ARG, "query " + std::to_string(i) + ": source out of range");
        if (!(qq.tp > 0.0 && qq.tp <= 1.0))
            return fail(ctx, PTW_ERR_INVALID_ARG, "query " + std::to_string(i) + ": tp must be in (0, 1]");
        if (qq.walkers == 0) return fail(ctx, PTW_ERR_INVALID_ARG, "query " + std::to_string(i) + ": no walkers");
        if (qq.kind != PTW_QUERY_VISITS && qq.kind != PTW_QUERY_ENDPOINTS)
            return fail(ctx, PTW_ERR_INVALID_ARG, "query " + std::to_string(i) + ": unknown kind");
        const bool ends = qq.kind == PTW_QUERY_ENDPOINTS;
        const double per_walker = batch_counts_per_walker(qq.tp, ends);
        // No u32 counter can pass the query's count total; plan for 1.5x the
        // expected total here and check the actual total after the walk.
        if ((double)qq.walkers * per_walker * 1.5 > 4294967296.0)
            return fail(ctx, PTW_ERR_INVALID_ARG,
                        "query " + std::to_string(i) + ": too many walkers for u32 slices; use ptw_walk");
        BatchQuery &B = P.q[i];
        B.source_dev = ctx->dev.relabelled ? ctx->h_perm[qq.source] : qq.source;
        B.thr = continue_threshold(qq.tp);
        B.walkers = qq.walkers;
        B.stream = qq.stream;
        B.ends_only = ends ? 1u : 0u;
        B.pad = 0;
        const SliceSizes z = batch_slice_sizes(V, qq.walkers, qq.tp, thr_dense, ends);
        if (z.dense && log_ok && (double)qq.walkers * per_walker <= kLogCountsPerVertex * V) {
            // no slice: a hub slice (below) and the log, whose pair list
            // the device places after the walk
            B.kind = SLICE_LOG;
            B.slice_off = 0;
            B.mask = 0;
            ++P.n_dense;
            P.log_entries += (uint64_t)(1.25 * (double)qq.walkers * per_walker) + 64;
            P.log_hubs += (uint64_t)log_hub_len + 1;
            P.log_order.push_back(i);
            P.log = true;
        } else if (z.dense) {
            B.kind = SLICE_DENSE;
            B.slice_off = P.dense;
            P.dense += z.slots;
            B.mask = 0;
            ++P.n_dense;
        } else {
            B.kind = SLICE_HASH;
            B.slice_off = P.hash;
            P.hash += z.slots;
            B.mask = (uint32_t)(z.slots - 1);
            ++P.n_hash;
        }
        // a multiple of 64 counters, like every dense slice: 256 B aligned
        B.hub_len = B.kind == SLICE_LOG ? log_hub_len : z.hub_len;
        B.hub_off = P.dense;
        P.dense += B.hub_len;
        B.touched_cap = z.touched_cap;
        B.touched_off = P.touched;
        if (B.kind != SLICE_LOG) P.touched += z.touched_cap;
        // top-k candidates: no more than the query's distinct vertices while
        // its pair list holds them all, and never more than the LDS tile
        B.cand_cap = d.topk ? std::min<uint32_t>((uint32_t)ctx->shape.topk_tile, std::max<uint32_t>(z.touched_cap, 1))
                            : 0;
        B.cand_off = P.cand;
        P.cand += B.cand_cap;
    }
    for (uint32_t i = 0; i < d.num_queries; ++i) {
        if (P.q[i].kind == SLICE_HASH) continue;
        const uint64_t dchunk =
            ctx->shape.dense_chunk ? (uint64_t)ctx->shape.dense_chunk : dense_chunk_size(P.q[i].walkers);
        for (uint64_t s = 0; s < P.q[i].walkers; s += dchunk)
            P.chunks.push_back(BatchChunk{i, (uint32_t)std::min<uint64_t>(dchunk, P.q[i].walkers - s), s});
    }
    P.dense_chunks = P.chunks.size();
    if (P.log) {
        // largest expected log first, so the last queries counted are small ones
        std::stable_sort(P.log_order.begin(), P.log_order.end(), [&](uint32_t x, uint32_t y) {
            const double cx = (double)d.queries[x].walkers * batch_counts_per_walker(d.queries[x].tp, P.q[x].ends_only != 0);
            const double cy = (double)d.queries[y].walkers * batch_counts_per_walker(d.queries[y].tp, P.q[y].ends_only != 0);
            return cx > cy;
        });
        P.log_segs = batch_log_segments(P.log_entries, P.dense_chunks, ctx->shape.block, ctx->num_cus);
        // visit_log = N > 1 caps the log at N entries (a full log reruns the batch with slices)
        if (ctx->shape.visit_log > 1)
            P.log_segs = std::min<uint64_t>(P.log_segs, (uint64_t)ctx->shape.visit_log / kBatchLogSeg);
    }
    for (uint32_t i = 0; i < d.num_queries; ++i) {
        if (P.q[i].kind != SLICE_HASH) continue;
        for (uint64_t s = 0; s < P.q[i].walkers; s += chunk)
            P.chunks.push_back(BatchChunk{i, (uint32_t)std::min<uint64_t>(chunk, P.q[i].walkers - s), s});
    }
    return PTW_OK;
}

uint64_t plan_bytes_beyond(const BatchState &b, const Plan &P, uint32_t nq, uint32_t k)
{
    uint64_t need = 0;
    if (P.dense > b.dense_cap) need += 4 * P.dense;
    if (P.hash > b.hash_cap) need += 8 * P.hash;
    if (P.touched_total() > b.touched_cap) need += 8 * P.touched_total();
    if (P.log) {
        if (P.log_segs > b.log_segs) need += P.log_segs * (kBatchLogSeg * 6 + 12);
        const uint64_t rows = (uint64_t)nq * P.log_nb;
        if (nq > b.log_q_cap) need += (uint64_t)nq * (4 * kBatchLogQWords + 8);
        if (rows > b.log_rows) need += rows * (12 + 8 * ((1ull << kBatchLogShift) / kBatchRunIds));
    }
    if (P.cand > b.cand_cap) need += 8 * P.cand;
    if (nq > b.q_cap) need += (uint64_t)nq * (sizeof(BatchQuery) + sizeof(TopkStats) + 24);
    if (P.chunks.size() > b.chunk_cap) need += P.chunks.size() * sizeof(BatchChunk);
    if ((uint64_t)nq * k > b.topk_cap) need += (uint64_t)nq * k * 12;
    return need;
}

// Kernel arguments hold addresses only; per-batch values live in hdr.
BatchLaunch launch_args(ptw_ctx *ctx, const BatchState &b)
{
    BatchLaunch p{};
    p.row_ptr = ctx->dev.row_ptr;
    p.col = ctx->dev.col;
    p.wedge = ctx->dev.wedge;
    p.nrec = ctx->dev.nrec;
    p.V = ctx->dev.V;
    p.queries = b.queries;
    p.chunks = b.chunks;
    p.hdr = b.hdr;
    p.dense = b.dense;
    p.hash = b.hash;
    p.touched = b.touched;
    p.touched_count = b.touched_count;
    p.overflow = b.overflow;
    p.steps = b.steps;
    p.entropy = b.entropy;
    p.queue = b.queue;
    p.tk_stats = b.tk_stats;
    p.tk_cand = b.tk_cand;
    if (b.vlog && b.log_active) {
        p.vlog = b.vlog;
        p.log_segs = b.log_segs;
        p.seg_q = b.seg_q;
        p.seg_n = b.seg_n;
        p.seg_list = b.seg_list;
        p.qlog = b.qlog;
        p.log_order = b.log_order;
        p.q_base = b.q_base;
        p.log_shift = b.log_shift;
        p.log_nb = b.log_nb;
        p.part_cnt = b.part_cnt;
        p.part_loc = b.part_loc;
        p.part_cur = b.part_cur;
        p.run_loc = b.run_loc;
        p.run_len = b.run_len;
        p.part16 = b.part16;
    }
    return p;
}

// Grows the pools to fit the plan; new slice memory is filled clean.
int ensure_pools(ptw_ctx *ctx, LaneRef L, const Plan &P, uint32_t nq, uint32_t k)
{
    BatchState &b = L.b;
    bool moved = false;
    if (!b.hdr) {
        HIP_TRY(ctx, hipMalloc((void **)&b.hdr, kBatchHdrWords * sizeof(unsigned long long)));
        HIP_TRY(ctx, hipMalloc((void **)&b.queue, 8 * sizeof(unsigned long long)));
        moved = true;
    }
    // the pools are kept when a batch has no logged queries (a free would wait
    // for the device); its launch does not run the log's kernels
    if (P.log != b.log_active) {
        b.log_active = P.log;
        moved = true;
    }
    if (P.log) {
        if (P.log_segs > b.log_segs) {
            hfree(b.vlog);
            hfree(b.seg_q);
            hfree(b.seg_n);
            hfree(b.seg_list);
            hfree(b.part16);
            const uint64_t cap = std::max<uint64_t>(P.log_segs, b.log_segs + b.log_segs / 2);
            HIP_TRY(ctx, hipMalloc((void **)&b.vlog, cap * kBatchLogSeg * 4));
            // 8 entries of padding: the counting reads aligned 16 B chunks
            HIP_TRY(ctx, hipMalloc((void **)&b.part16, cap * kBatchLogSeg * 2 + 16));
            HIP_TRY(ctx, hipMalloc((void **)&b.seg_q, cap * 4));
            HIP_TRY(ctx, hipMalloc((void **)&b.seg_n, cap * 4));
            HIP_TRY(ctx, hipMalloc((void **)&b.seg_list, cap * 4));
            b.log_segs = cap;
            moved = true;
        }
        if (nq > b.log_q_cap || !b.qlog) {
            hfree(b.qlog);
            hfree(b.q_base);
            hfree(b.log_order);
            const uint64_t cap = std::max<uint64_t>(std::max<uint64_t>(nq, 1), b.log_q_cap + b.log_q_cap / 2);
            HIP_TRY(ctx, hipMalloc((void **)&b.qlog, cap * kBatchLogQWords * 4));
            HIP_TRY(ctx, hipMalloc((void **)&b.log_order, cap * 4));
            HIP_TRY(ctx, hipMalloc((void **)&b.q_base, cap * 8));
            b.log_q_cap = cap;
            moved = true;
        }
        const uint64_t rows = std::max<uint64_t>((uint64_t)nq * P.log_nb, 1);
        if (rows > b.log_rows || !b.part_cnt) {
            hfree(b.part_cnt);
            hfree(b.part_loc);
            hfree(b.part_cur);
            hfree(b.run_loc);
            hfree(b.run_len);
            const uint64_t cap = std::max<uint64_t>(rows, b.log_rows + b.log_rows / 2);
            HIP_TRY(ctx, hipMalloc((void **)&b.part_cnt, cap * 4));
            HIP_TRY(ctx, hipMalloc((void **)&b.part_loc, cap * 4));
            HIP_TRY(ctx, hipMalloc((void **)&b.part_cur, cap * 4));
            const uint64_t runs = cap * ((1ull << kBatchLogShift) / kBatchRunIds);
            HIP_TRY(ctx, hipMalloc((void **)&b.run_loc, runs * 4));
            HIP_TRY(ctx, hipMalloc((void **)&b.run_len, runs * 4));
            b.log_rows = cap;
            moved = true;
        }
        if (b.log_shift != P.log_shift || b.log_nb != P.log_nb) {
            b.log_shift = P.log_shift;
            b.log_nb = P.log_nb;
            moved = true;
        }
    }
    if (P.dense > b.dense_cap) {
        hfree(b.dense);
        const uint64_t cap = std::max<uint64_t>(P.dense, b.dense_cap + b.dense_cap / 2);
        HIP_TRY(ctx, hipMalloc((void **)&b.dense, cap * 4));
        HIP_TRY(ctx, launch_fill_u32(b.dense, cap, 0u, L.s));
        b.dense_cap = cap;
        moved = true;
    }
    if (P.hash > b.hash_cap) {
        hfree(b.hash);
        const uint64_t cap = std::max<uint64_t>(P.hash, b.hash_cap + b.hash_cap / 2);
        HIP_TRY(ctx, hipMalloc((void **)&b.hash, cap * 8));
        // {key = empty, count = 0}: 0x00000000FFFFFFFF little-endian per slot
        HIP_TRY(ctx, launch_fill_u64((unsigned long long *)b.hash, cap, 0x00000000FFFFFFFFull, L.s));
        b.hash_cap = cap;
        moved = true;
    }
    if (P.touched_total() > b.touched_cap || !b.touched) {
        hfree(b.touched);
        const uint64_t cap =
            std::max<uint64_t>(std::max<uint64_t>(P.touched_total(), 1), b.touched_cap + b.touched_cap / 2);
        HIP_TRY(ctx, hipMalloc((void **)&b.touched, cap * 8));
        b.touched_cap = cap;
        moved = true;
    }
    if (nq > b.q_cap) {
        hfree(b.queries);
        hfree(b.touched_count);
        hfree(b.overflow);
        hfree(b.steps);
        hfree(b.entropy);
        hfree(b.tk_stats);
        const uint32_t cap = std::max<uint32_t>(nq, b.q_cap + b.q_cap / 2);
        HIP_TRY(ctx, hipMalloc((void **)&b.queries, (size_t)cap * sizeof(BatchQuery)));
        HIP_TRY(ctx, hipMalloc((void **)&b.touched_count, (size_t)cap * 4));
        HIP_TRY(ctx, hipMalloc((void **)&b.overflow, (size_t)cap * 4));
        HIP_TRY(ctx, hipMalloc((void **)&b.steps, (size_t)cap * 8));
        HIP_TRY(ctx, hipMalloc((void **)&b.entropy, (size_t)cap * 8));
        HIP_TRY(ctx, hipMalloc((void **)&b.tk_stats, (size_t)cap * sizeof(TopkStats)));
        b.q_cap = cap;
        moved = true;
    }
    if (P.cand > b.cand_cap || !b.tk_cand) {
        hfree(b.tk_cand);
        const uint64_t cap = std::max<uint64_t>(std::max<uint64_t>(P.cand, 1), b.cand_cap + b.cand_cap / 2);
        HIP_TRY(ctx, hipMalloc((void **)&b.tk_cand, cap * 8));
        b.cand_cap = cap;
        moved = true;
    }
    if (P.chunks.size() > b.chunk_cap) {
        hfree(b.chunks);
        const uint64_t cap = std::max<uint64_t>(P.chunks.size(), b.chunk_cap + b.chunk_cap / 2);
        HIP_TRY(ctx, hipMalloc((void **)&b.chunks, cap * sizeof(BatchChunk)));
        b.chunk_cap = cap;
        moved = true;
    }
    const uint64_t tk = std::max<uint64_t>((uint64_t)nq * k, 1);
    if (tk > b.topk_cap) {
        hfree(b.topk_v);
        hfree(b.topk_c);
        const uint64_t cap = std::max<uint64_t>(tk, b.topk_cap + b.topk_cap / 2);
        HIP_TRY(ctx, hipMalloc((void **)&b.topk_v, cap * 4));
        HIP_TRY(ctx, hipMalloc((void **)&b.topk_c, cap * 8));
        b.topk_cap = cap;
        moved = true;
    }
    if (moved) drop_exec(b);
    return PTW_OK;
}

// Counter setup + walk + top-k for the batch described by the device tables.
int enqueue_run(ptw_ctx *ctx, LaneRef L, const BatchLaunch &p)
{
    HIP_TRY(ctx, launch_batch_begin(p, L.b.q_cap, L.s));
    HIP_TRY(ctx, launch_batch_walk(p, ctx->shape.block, ctx->num_cus, L.cus, ctx->shape.lds_hubs,
                                   (int)ptw::batch_cache_slots(ctx->shape), L.s));
    HIP_TRY(ctx, launch_batch_topk(p, ctx->dev.inv, L.b.topk_v, L.b.topk_c, ctx->shape.topk_tile, ctx->num_cus,
                                   L.s));
    return PTW_OK;
}

int clean_previous(ptw_ctx *ctx, LaneRef L)
{
    BatchState &b = L.b;
    if (!b.dirty) return PTW_OK;
    BatchLaunch p = launch_args(ctx, b);
    HIP_TRY(ctx, launch_batch_reset(p, ctx->num_cus, L.s));
    b.dirty = false;
    return PTW_OK;
}

void free_state(BatchState &b)
{
    drop_exec(b);
    free_log(b);
    hfree(b.dense);
    hfree(b.hash);
    hfree(b.touched);
    hfree(b.queries);
    hfree(b.touched_count);
    hfree(b.overflow);
    hfree(b.steps);
    hfree(b.entropy);
    hfree(b.tk_stats);
    hfree(b.tk_cand);
    hfree(b.chunks);
    hfree(b.topk_v);
    hfree(b.topk_c);
    hfree(b.hdr);
    hfree(b.queue);
    hfree(b.g_q);
    hfree(b.g_v);
    hfree(b.g_out);
    b = BatchState{};
}

}  // namespace

int batch_enqueue(ptw_ctx *ctx, LaneRef L, const ptw_batch_desc &d, bool allow_log)
{
    if (!ctx->has_graph) return fail(ctx, PTW_ERR_NO_GRAPH, "no graph loaded");
    if (ctx->device < 0) return fail(ctx, PTW_ERR_UNSUPPORTED, "host-only context: batches need a HIP device");
    if (d.num_queries && !d.queries) return fail(ctx, PTW_ERR_INVALID_ARG, "null query array");
    if (d.topk > 1024 || 2 * d.topk > (uint32_t)ctx->shape.topk_tile)
        return fail(ctx, PTW_ERR_INVALID_ARG, "topk must be <= 1024 and <= topk_tile / 2");
    if (d.semantics != PTW_SEM_REFERENCE && d.semantics != PTW_SEM_TEXTBOOK)
        return fail(ctx, PTW_ERR_INVALID_ARG, "unknown semantics");
    const bool weighted = (d.flags & PTW_BATCH_WEIGHTED) != 0;
    if (weighted && !ctx->dev.wedge) return fail(ctx, PTW_ERR_INVALID_ARG, "weighted batch on a graph without weights");
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    BatchState &b = L.b;
    Plan P;
    int rc = make_plan(ctx, d, P, allow_log);
    if (rc) return rc;
    // admission: what the plan needs beyond the pools must fit in free HBM
    const uint64_t extra = plan_bytes_beyond(b, P, d.num_queries, d.topk);
    if (extra) {
        size_t f = 0, t = 0;
        HIP_TRY(ctx, hipMemGetInfo(&f, &t));
        if (extra > f)
            return fail(ctx, PTW_ERR_ADMISSION,
                        "batch needs " + std::to_string(extra) + " more bytes of HBM, " + std::to_string(f) + " free");
    }
    if ((rc = clean_previous(ctx, L))) return rc;
    if ((rc = ensure_pools(ctx, L, P, d.num_queries, d.topk))) return rc;
    // neighbour records (built once; a captured launch without them is redone)
    if (!weighted && ctx->shape.edge_records == 1 && !ctx->dev.nrec && ensure_edge_records(ctx, false))
        drop_exec(b);

    b.h_queries = std::move(P.q);
    b.h_chunks = std::move(P.chunks);
    b.h_hdr[0] = b.h_chunks.size();
    b.h_hdr[1] = d.num_queries;
    b.h_hdr[2] = d.topk;
    b.h_hdr[3] = d.seed;
    b.h_hdr[4] = (unsigned long long)d.semantics;
    b.h_hdr[5] = weighted ? 1 : 0;
    b.h_hdr[6] = P.dense_chunks;
    b.h_hdr[7] = P.touched;  // logged queries' pair lists follow the others
    b.h_hdr[8] = P.log_segs;
    b.h_hdr[9] = P.log_order.size();
    b.h_log_order = std::move(P.log_order);
    b.last_nq = d.num_queries;
    b.weighted = weighted;
    hipStream_t s = L.s;
    if (d.num_queries) {
        HIP_TRY(ctx, hipMemcpyAsync(b.queries, b.h_queries.data(), b.h_queries.size() * sizeof(BatchQuery),
                                    hipMemcpyHostToDevice, s));
        HIP_TRY(ctx, hipMemcpyAsync(b.chunks, b.h_chunks.data(), b.h_chunks.size() * sizeof(BatchChunk),
                                    hipMemcpyHostToDevice, s));
        if (!b.h_log_order.empty())
            HIP_TRY(ctx, hipMemcpyAsync(b.log_order, b.h_log_order.data(), b.h_log_order.size() * 4,
                                        hipMemcpyHostToDevice, s));
    }
    HIP_TRY(ctx, hipMemcpyAsync(b.hdr, b.h_hdr, sizeof(b.h_hdr), hipMemcpyHostToDevice, s));
    b.dirty = true;

    const BatchLaunch p = launch_args(ctx, b);
    HIP_TRY(ctx, hipEventRecord(L.ev0, s));
    if (d.flags & PTW_BATCH_GRAPH) {
        // The sequence reads sizes, seed and semantics from hdr, so one capture
        // serves every batch until a pool or the graph moves.
        if (!b.exec) {
            hipGraph_t g = nullptr;
            HIP_TRY(ctx, hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
            rc = enqueue_run(ctx, L, p);
            hipError_t ec = hipStreamEndCapture(s, &g);
            if (rc) {
                if (g) (void)hipGraphDestroy(g);
                return rc;
            }
            HIP_TRY(ctx, ec);
            hipError_t ei = hipGraphInstantiate(&b.exec, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            HIP_TRY(ctx, ei);
        }
        HIP_TRY(ctx, hipGraphLaunch(b.exec, s));
    } else {
        if ((rc = enqueue_run(ctx, L, p))) return rc;
    }
    HIP_TRY(ctx, hipEventRecord(L.ev1, s));
    return PTW_OK;
}

int batch_enqueue_results(ptw_ctx *ctx, LaneRef L, uint32_t nq, uint32_t k, unsigned long long *steps,
                          uint32_t *over, double *ent, uint32_t *tv, unsigned long long *tc)
{
    if (!nq) return PTW_OK;
    const BatchState &b = L.b;
    if (steps) HIP_TRY(ctx, hipMemcpyAsync(steps, b.steps, (size_t)nq * 8, hipMemcpyDeviceToHost, L.s));
    if (over) HIP_TRY(ctx, hipMemcpyAsync(over, b.overflow, (size_t)nq * 4, hipMemcpyDeviceToHost, L.s));
    if (ent) HIP_TRY(ctx, hipMemcpyAsync(ent, b.entropy, (size_t)nq * 8, hipMemcpyDeviceToHost, L.s));
    if (k && tv) HIP_TRY(ctx, hipMemcpyAsync(tv, b.topk_v, (size_t)nq * k * 4, hipMemcpyDeviceToHost, L.s));
    if (k && tc) HIP_TRY(ctx, hipMemcpyAsync(tc, b.topk_c, (size_t)nq * k * 8, hipMemcpyDeviceToHost, L.s));
    return PTW_OK;
}

int batch_check(ptw_ctx *ctx, const ptw_batch_desc &d, const unsigned long long *steps, const uint32_t *over)
{
    for (uint32_t i = 0; i < d.num_queries; ++i) {
        if (over[i])
            return fail(ctx, PTW_ERR_OOM, "query " + std::to_string(i) + " overflowed its hash slice or the visit log");
        const uint64_t total_counts =
            d.queries[i].walkers + (d.queries[i].kind == PTW_QUERY_ENDPOINTS ? 0 : steps[i]);
        if (total_counts >= 4294967296ull)
            return fail(ctx, PTW_ERR_OOM,
                        "query " + std::to_string(i) + ": count total reached 2^32, u32 counters may have wrapped");
    }
    return PTW_OK;
}

int batch_gather_lane(ptw_ctx *ctx, LaneRef L, uint64_t n, const uint32_t *query_index, const uint32_t *vertices,
                      uint64_t *out)
{
    BatchState &b = L.b;
    if (!b.dirty) return fail(ctx, PTW_ERR_INVALID_ARG, "no batch results to gather from");
    for (uint64_t i = 0; i < n; ++i)
        if (query_index[i] >= b.last_nq || vertices[i] >= ctx->dev.V)
            return fail(ctx, PTW_ERR_INVALID_ARG, "gather pair " + std::to_string(i) + " out of range");
    if (n == 0) return PTW_OK;
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    hipStream_t s = L.s;
    // grown and kept: a gather frees nothing, so it never waits for the work
    // of other streams (a free does)
    if (n > b.g_cap) {
        HIP_TRY(ctx, hipStreamSynchronize(s));
        hfree(b.g_q);
        hfree(b.g_v);
        hfree(b.g_out);
        b.g_cap = 0;
        const uint64_t cap = std::max<uint64_t>(n, 4096);
        HIP_TRY(ctx, hipMalloc((void **)&b.g_q, cap * 4));
        HIP_TRY(ctx, hipMalloc((void **)&b.g_v, cap * 4));
        HIP_TRY(ctx, hipMalloc((void **)&b.g_out, cap * 8));
        b.g_cap = cap;
    }
    HIP_TRY(ctx, hipMemcpyAsync(b.g_q, query_index, n * 4, hipMemcpyHostToDevice, s));
    HIP_TRY(ctx, hipMemcpyAsync(b.g_v, vertices, n * 4, hipMemcpyHostToDevice, s));
    const BatchLaunch p = launch_args(ctx, b);
    HIP_TRY(ctx, launch_batch_gather(p, n, b.g_q, b.g_v, ctx->dev.perm, b.g_out, s));
    HIP_TRY(ctx, hipMemcpyAsync(out, b.g_out, n * 8, hipMemcpyDeviceToHost, s));
    HIP_TRY(ctx, hipStreamSynchronize(s));
    return PTW_OK;
}

void free_batch(ptw_ctx *ctx)
{
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    free_state(ctx->batch);
    for (auto &l : ctx->serve) {
        if (l.stream) (void)hipStreamSynchronize(l.stream);
        free_state(l.b);
    }
}

void destroy_serve_lanes(ptw_ctx *ctx)
{
    for (auto &l : ctx->serve) {
        if (l.stream) (void)hipStreamSynchronize(l.stream);
        free_state(l.b);
        if (l.ev0) (void)hipEventDestroy(l.ev0);
        if (l.ev1) (void)hipEventDestroy(l.ev1);
        if (l.done) (void)hipEventDestroy(l.done);
        if (l.stream) (void)hipStreamDestroy(l.stream);
        if (l.h_steps) (void)hipHostFree(l.h_steps);
        if (l.h_over) (void)hipHostFree(l.h_over);
        if (l.h_ent) (void)hipHostFree(l.h_ent);
        if (l.h_tv) (void)hipHostFree(l.h_tv);
        if (l.h_tc) (void)hipHostFree(l.h_tc);
        l = ServeLane{};
    }
}

}  // namespace ptw

using ptw::fail;

extern "C" {

int ptw_batch_plan_bytes(ptw_ctx *ctx, const ptw_batch_desc *desc, uint64_t *bytes)
{
    if (!ctx || !desc || !bytes) return fail(ctx, PTW_ERR_INVALID_ARG, "null argument");
    if (!ctx->has_graph) return fail(ctx, PTW_ERR_NO_GRAPH, "no graph loaded");
    if (desc->num_queries && !desc->queries) return fail(ctx, PTW_ERR_INVALID_ARG, "null query array");
    ptw::Plan P;
    int rc = ptw::make_plan(ctx, *desc, P);
    if (rc) return rc;
    *bytes = ptw::plan_bytes_beyond(ctx->batch, P, desc->num_queries, desc->topk);
    return PTW_OK;
}

int ptw_batch_ppr(ptw_ctx *ctx, const ptw_batch_desc *desc, ptw_batch_out *out)
{
    if (!ctx) return PTW_ERR_INVALID_ARG;
    if (!desc || !out) return fail(ctx, PTW_ERR_INVALID_ARG, "null descriptor or output");
    if (!ctx->has_graph) return fail(ctx, PTW_ERR_NO_GRAPH, "no graph loaded");
    if (ctx->device < 0) return fail(ctx, PTW_ERR_UNSUPPORTED, "host-only context: batches need a HIP device");
    const ptw_batch_desc &d = *desc;
    if (d.num_queries && !d.queries) return fail(ctx, PTW_ERR_INVALID_ARG, "null query array");
    if (d.topk > 1024 || 2 * d.topk > (uint32_t)ctx->shape.topk_tile)
        return fail(ctx, PTW_ERR_INVALID_ARG, "topk must be <= 1024 and <= topk_tile / 2");
    if (d.topk && (!out->topk_vertex || !out->topk_count))
        return fail(ctx, PTW_ERR_INVALID_ARG, "top-k output arrays required");
    const double t0 = ptw::now_ms();
    out->total_steps = 0;
    out->slice_bytes = 0;
    out->dense_queries = out->hash_queries = 0;
    out->logged_queries = out->log_rerun = 0;
    out->kernel_ms = 0.0;
    const ptw::LaneRef L = ptw::main_lane(ctx);
    std::vector<unsigned long long> steps(d.num_queries), tc((size_t)d.num_queries * d.topk);
    std::vector<uint32_t> over(d.num_queries);
    auto run = [&](bool allow_log) -> int {
        int r = ptw::batch_enqueue(ctx, L, d, allow_log);
        if (r) return r;
        if ((r = ptw::batch_enqueue_results(ctx, L, d.num_queries, d.topk, steps.data(), over.data(), out->entropy,
                                            d.topk ? out->topk_vertex : nullptr, tc.data())))
            return r;
        HIP_TRY(ctx, hipStreamSynchronize(L.s));
        return PTW_OK;
    };
    int rc = run(true);
    if (rc) return rc;
    bool logged = false;
    for (const auto &q : L.b.h_queries) logged |= q.kind == ptw::SLICE_LOG;
    if (logged && std::any_of(over.begin(), over.end(), [](uint32_t o) { return o != 0; })) {
        // the visit log filled: the same batch with slices gives the same counts
        if ((rc = run(false))) return rc;
        out->log_rerun = 1;
    }
    float ms = 0.f;
    HIP_TRY(ctx, hipEventElapsedTime(&ms, L.ev0, L.ev1));
    out->kernel_ms = ms;
    for (size_t i = 0; i < tc.size(); ++i) out->topk_count[i] = tc[i];
    uint64_t total = 0;
    for (uint32_t i = 0; i < d.num_queries; ++i) {
        total += steps[i];
        if (out->steps) out->steps[i] = steps[i];
    }
    out->total_steps = total;
    uint64_t bytes = 0;
    for (const auto &q : L.b.h_queries) {
        if (q.kind == ptw::SLICE_DENSE) {
            ++out->dense_queries;
            bytes += 4ull * ptw::dense_pitch(ctx->dev.V);
        } else if (q.kind == ptw::SLICE_LOG) {
            ++out->dense_queries;
            ++out->logged_queries;
            const ptw_query &qq = d.queries[&q - L.b.h_queries.data()];
            bytes += ptw::batch_query_bytes(ctx->dev.V, qq.walkers, qq.tp, qq.kind == PTW_QUERY_ENDPOINTS, 0, true,
                                            q.hub_len);
            continue;
        } else {
            ++out->hash_queries;
            bytes += 8ull * ((uint64_t)q.mask + 1) + 4ull * q.hub_len;
        }
        bytes += 8ull * q.touched_cap;
    }
    out->slice_bytes = bytes;
    out->total_ms = ptw::now_ms() - t0;
    return ptw::batch_check(ctx, d, steps.data(), over.data());
}

int ptw_batch_write_topk(ptw_ctx *ctx, const char *path, const ptw_batch_desc *desc, const ptw_batch_out *out,
                         int id_space)
{
    if (!ctx || !path || !desc || !out) return fail(ctx, PTW_ERR_INVALID_ARG, "null argument");
    if (!ctx->has_graph) return fail(ctx, PTW_ERR_NO_GRAPH, "no graph loaded");
    if (id_space != PTW_IDS_GRAPH && id_space != PTW_IDS_ORIGINAL)
        return fail(ctx, PTW_ERR_INVALID_ARG, "unknown id space");
    const uint32_t n = desc->num_queries, k = desc->topk;
    if ((n && !desc->queries) || (n && k && (!out->topk_vertex || !out->topk_count)))
        return fail(ctx, PTW_ERR_INVALID_ARG, "null query or top-k array");
    const ptw::HostCsr &h = ctx->host;
    auto id = [&](uint32_t v) { return id_space == PTW_IDS_ORIGINAL ? h.name(v) : v; };
    FILE *f = fopen(path, "w");
    if (!f) return fail(ctx, PTW_ERR_IO, std::string("cannot open '") + path + "' for writing");
    bool ok = fprintf(f, "# query source kind rank vertex count\n") > 0;
    for (uint32_t q = 0; q < n && ok; ++q) {
        const ptw_query &qq = desc->queries[q];
        if (qq.source >= h.V) {
            fclose(f);
            return fail(ctx, PTW_ERR_INVALID_ARG, "query " + std::to_string(q) + ": source out of range");
        }
        for (uint32_t r = 0; r < k && ok; ++r) {
            const uint32_t v = out->topk_vertex[(uint64_t)q * k + r];
            if (v == PTW_INVALID_VERTEX) break;
            if (v >= h.V) {
                fclose(f);
                return fail(ctx, PTW_ERR_INVALID_ARG, "top-k entry out of range");
            }
            ok = fprintf(f, "%u %u %u %u %u %llu\n", q, id(qq.source), qq.kind, r, id(v),
                         (unsigned long long)out->topk_count[(uint64_t)q * k + r]) > 0;
        }
    }
    if (fclose(f) != 0) ok = false;
    if (!ok) return fail(ctx, PTW_ERR_IO, std::string("short write to '") + path + "'");
    return PTW_OK;
}

int ptw_batch_gather(ptw_ctx *ctx, uint64_t n, const uint32_t *query_index, const uint32_t *vertices,
                     uint64_t *out)
{
    if (!ctx || (n && (!query_index || !vertices || !out))) return fail(ctx, PTW_ERR_INVALID_ARG, "null argument");
    if (ctx->device < 0) return fail(ctx, PTW_ERR_UNSUPPORTED, "host-only context");
    return ptw::batch_gather_lane(ctx, ptw::main_lane(ctx), n, query_index, vertices, out);
}

int ptw_batch_counts(ptw_ctx *ctx, uint32_t index, uint64_t *counts)
{
    if (!ctx || !counts) return fail(ctx, PTW_ERR_INVALID_ARG, "null argument");
    if (ctx->device < 0) return fail(ctx, PTW_ERR_UNSUPPORTED, "host-only context");
    ptw::BatchState &b = ctx->batch;
    if (!b.dirty || index >= b.last_nq) return fail(ctx, PTW_ERR_INVALID_ARG, "no such query in the last batch");
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    const uint32_t V = ctx->dev.V;
    uint64_t *tmp = nullptr;
    HIP_TRY(ctx, hipMalloc((void **)&tmp, (size_t)V * 8));
    const ptw::BatchLaunch p = ptw::launch_args(ctx, b);
    hipError_t e = ptw::launch_batch_export(p, index, ctx->dev.perm, ctx->dev.inv, tmp, ctx->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(counts, tmp, (size_t)V * 8, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    (void)hipFree(tmp);
    HIP_TRY(ctx, e);
    return PTW_OK;
}

}  // extern "C"
