// C ABI of the walk engine (include/ptw.h).  Host orchestration only: the
// walk itself runs in the HIP kernels of walk.hip; there is no CPU path.
#include <hip/hip_runtime.h>
#include <sys/stat.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>
#include <exception>
#include <new>
#include <string>

#include "../../include/ptw.h"
#include "ctx.hpp"
#include "prims.hpp"
#include "engine.hpp"

namespace {

using ptw::fail;
using ptw::HostCsr;
using ptw::upload;

bool is_count_app(int app) { return app == PTW_APP_PAGERANK || app == PTW_APP_PPR; }
bool is_path_app(int app) { return app == PTW_APP_DEEPWALK || app == PTW_APP_NODE2VEC; }

void free_vlog(ptw_ctx *ctx)
{
    ptw::VlogBuffers &b = ctx->vlog;
    if (b.log) (void)hipFree(b.log);
    if (b.parted) (void)hipFree(b.parted);
    if (b.plan) (void)hipFree(b.plan);
    b = ptw::VlogBuffers{};
}

// The visit log for a count walk of `walkers` walkers at termination
// probability tp: room for 1.1x the expected visits plus one segment per
// resident wavefront (each may leave one partly used), at most
// kVlogMaxEntries and at most `limit` entries (a multiple of kVlogSeg; 0:
// no limit).  The buffers grow and are kept.  Returns the entries the walk
// may use, or 0 (device atomics) if the graph needs too many buckets or the
// memory is not there.
uint64_t ensure_vlog(ptw_ctx *ctx, uint64_t walkers, double tp, uint64_t limit = 0)
{
    const uint32_t V = ctx->dev.V;
    const uint32_t shift = ptw::vlog_shift_for(V);
    if (!shift) return 0;
    const uint32_t nb = ptw::vlog_buckets_for(V, shift);
    const double visits = 1.1 * (double)walkers * (1.0 + 1.0 / tp) + 64.0 * ctx->num_cus * ptw::kVlogSeg;
    uint64_t cap = visits >= (double)ptw::kVlogMaxEntries ? ptw::kVlogMaxEntries : (uint64_t)visits;
    if (limit && cap > limit) cap = limit;
    cap = (cap + ptw::kVlogSeg - 1) / ptw::kVlogSeg * ptw::kVlogSeg;
    ptw::VlogBuffers &b = ctx->vlog;
    if (b.cap >= cap && b.plan_buckets >= nb) return cap;
    const uint64_t alloc = std::max(cap, b.cap);
    free_vlog(ctx);
    if (hipMalloc((void **)&b.log, alloc * 4) != hipSuccess || hipMalloc((void **)&b.parted, alloc * 2) != hipSuccess ||
        hipMalloc((void **)&b.plan, ptw::vlog_plan_words(nb) * 4) != hipSuccess) {
        (void)hipGetLastError();  // the failed allocation is not a walk error
        free_vlog(ctx);
        return 0;
    }
    b.cap = alloc;
    b.plan_buckets = nb;
    return cap;
}

}  // namespace

// Neighbour records for walks (built on first use, kept with the graph):
// unweighted {neighbour, row start, degree, graph id} per slot, or the
// weighted pair per slot for path walks.  False (the walk reads the CSR) if
// the memory is not there.  The build is waited for, so walks and batches on
// other streams can read the records at once.
bool ptw::ensure_edge_records(ptw_ctx *ctx, bool weighted)
{
    ptw::DeviceGraph &d = ctx->dev;
    if (weighted ? d.wnrec != nullptr : d.nrec != nullptr) return true;
    if (weighted && !d.wedge) return false;
    const uint64_t n = (weighted ? 2 : 1) * std::max<uint64_t>(d.E, 1);
    uint4 *rec = nullptr;
    if (hipMalloc((void **)&rec, n * sizeof(uint4)) != hipSuccess) {
        (void)hipGetLastError();  // a missing table is not a walk error
        return false;
    }
    const hipError_t e = weighted ? ptw::launch_weighted_records(d.row_ptr, d.wedge, d.inv, d.E, rec, ctx->stream)
                                  : ptw::launch_edge_records(d.row_ptr, d.col, d.inv, d.E, rec, ctx->stream);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        (void)hipFree(rec);
        return false;
    }
    if (hipStreamSynchronize(ctx->stream) != hipSuccess) {
        (void)hipGetLastError();
        (void)hipFree(rec);
        return false;
    }
    (weighted ? d.wnrec : d.nrec) = rec;
    d.bytes += n * sizeof(uint4);
    return true;
}

namespace {

void free_edge_records(ptw_ctx *ctx, bool weighted_only)
{
    ptw::DeviceGraph &d = ctx->dev;
    if (d.wnrec) {
        (void)hipFree(d.wnrec);
        d.bytes -= 32ull * std::max<uint64_t>(d.E, 1);
        d.wnrec = nullptr;
    }
    if (!weighted_only && d.nrec) {
        (void)hipFree(d.nrec);
        d.bytes -= 16ull * std::max<uint64_t>(d.E, 1);
        d.nrec = nullptr;
    }
}

// Which records a walk reads (setting: -1 path walks, 0 none, 1 every
// walk).  Weighted records serve path walks only; other weighted walks read
// the weighted-edge records.
void pick_edge_records(ptw_ctx *ctx, const ptw_walk_desc &d, bool may_build, ptw::WalkLaunch &p)
{
    const int setting = ctx->shape.edge_records;
    if (setting == 0) return;
    if (setting < 0 && !is_path_app(d.app)) return;
    const bool weighted = (d.flags & PTW_FLAG_WEIGHTED) != 0;
    if (weighted && !is_path_app(d.app)) return;
    if (!weighted ? ctx->dev.nrec == nullptr : ctx->dev.wnrec == nullptr) {
        if (!may_build || !ptw::ensure_edge_records(ctx, weighted)) return;
    }
    if (weighted)
        p.wnrec = ctx->dev.wnrec;
    else
        p.nrec = ctx->dev.nrec;
}

void free_graph(ptw_ctx *ctx)
{
    ptw::free_batch(ctx);
    free_vlog(ctx);
    free_edge_records(ctx, false);
    auto &d = ctx->dev;
    if (d.row_ptr) (void)hipFree(d.row_ptr);
    if (d.col) (void)hipFree(d.col);
    if (d.perm) (void)hipFree(d.perm);
    if (d.inv) (void)hipFree(d.inv);
    if (d.wedge) (void)hipFree(d.wedge);
    if (d.t_col) (void)hipFree(d.t_col);
    if (d.eset) (void)hipFree(d.eset);
    if (d.out_deg) (void)hipFree(d.out_deg);
    if (d.piece_begin) (void)hipFree(d.piece_begin);
    if (d.vertex_piece) (void)hipFree(d.vertex_piece);
    if (ctx->counts_dev) (void)hipFree(ctx->counts_dev);
    if (ctx->counts_stage) (void)hipFree(ctx->counts_stage);
    ctx->counts_stage = nullptr;
    if (ctx->pinned) (void)hipHostFree(ctx->pinned);
    ctx->pinned = nullptr;
    d = ptw::DeviceGraph{};
    ctx->counts_dev = nullptr;
    ctx->has_graph = false;
    ctx->host = HostCsr{};
    ctx->h_perm.clear();
    ctx->h_inv.clear();
    ctx->h_drow.clear();
    ctx->weights.clear();
    ctx->alias_thr.clear();
    ctx->alias_idx.clear();
}

// Builds alias tables from ctx->weights and uploads them in device order as
// weighted-edge records {neighbour, threshold, alias slot's neighbour, 0}
// (device ids), so a weighted pick is one 16 B read.
int install_weights(ptw_ctx *ctx)
{
    int rc = ptw::build_alias_tables(ctx->host, ctx->weights, ctx->alias_thr, ctx->alias_idx, ctx->err);
    if (rc) return rc;
    if (ctx->device < 0) return PTW_OK;
    auto &d = ctx->dev;
    const HostCsr &h = ctx->host;
    std::vector<uint4> tab(h.E());
    const bool rel = d.relabelled;
    ptw::parallel_for(h.V, [&](uint64_t lo, uint64_t hi) {
        for (uint64_t dv = lo; dv < hi; ++dv) {
            const uint32_t u = rel ? ctx->h_inv[dv] : (uint32_t)dv;
            const uint32_t src = h.row_ptr[u], n = h.row_ptr[u + 1] - src, dst = rel ? ctx->h_drow[dv] : src;
            auto nbr = [&](uint32_t k) { return rel ? ctx->h_perm[h.col[src + k]] : h.col[src + k]; };
            for (uint32_t k = 0; k < n; ++k)
                tab[dst + k] = uint4{nbr(k), ctx->alias_thr[src + k], nbr(ctx->alias_idx[src + k]), 0u};
        }
    });
    ptw::drop_batch_graph(ctx);
    free_edge_records(ctx, true);  // the weighted records follow the weights
    if (d.wedge) {
        (void)hipFree(d.wedge);
        d.bytes -= 16ull * h.E();
    }
    d.wedge = nullptr;
    if ((rc = upload(ctx, &d.wedge, tab.data(), tab.size()))) return rc;
    d.bytes += 16ull * h.E();
    return PTW_OK;
}

// Takes ownership of `g` (graph ids), builds the device copy.
// The first launch of each walk kernel and the first device-to-pageable copy
// of a process carry one-time runtime setup (several ms).  They are taken
// here, with the graph, so that no walk pays for them: zero-walker launches
// exit at their first queue take and touch nothing but the queue word.
int warm_up(ptw_ctx *ctx)
{
    hipStream_t s = ctx->stream;
    ptw::WalkLaunch p{};
    p.row_ptr = ctx->dev.row_ptr;
    p.col = ctx->dev.col;
    p.perm = ctx->dev.perm;
    p.inv = ctx->dev.inv;
    p.V = ctx->dev.V;
    p.hubs = ctx->dev.hubs;
    p.start_mode = PTW_START_SOURCE;
    p.walk_len = 1;
    p.walkers = 0;
    p.counts_dev = ctx->counts_dev;
    p.scratch = ctx->scratch;
    for (int app : {PTW_APP_PAGERANK, PTW_APP_RWR, PTW_APP_DEEPWALK, PTW_APP_NODE2VEC}) {
        p.app = app;
        HIP_TRY(ctx, ptw::launch_fill_u64(ctx->scratch, 8, 0ull, s));
        HIP_TRY(ctx, ptw::launch_walk(p, ctx->shape, ctx->num_cus, s));
    }
    // neighbour records belong to the graph's set-up, like the CSR: walks
    // do not pay for building them (a failed allocation leaves walks on the CSR)
    if (ctx->shape.edge_records != 0) (void)ptw::ensure_edge_records(ctx, false);
    // the visit log's passes over an empty log (a minimal log; count walks
    // grow it to their size)
    if (const uint64_t cap = ensure_vlog(ctx, 0, 1.0, ptw::kVlogSeg)) {
        const uint32_t shift = ptw::vlog_shift_for(ctx->dev.V);
        const uint32_t nb = ptw::vlog_buckets_for(ctx->dev.V, shift);
        HIP_TRY(ctx, ptw::launch_vlog_begin(ctx->vlog, nb, s));
        HIP_TRY(ctx, ptw::launch_vlog_count(ctx->vlog, ctx->scratch + 3, cap, shift, nb, ctx->counts_dev,
                                            ctx->num_cus, s));
    }
    HIP_TRY(ctx, ptw::launch_counts_to_graph_ids(ctx->counts_dev, ctx->dev.perm, std::min<uint32_t>(ctx->dev.V, 64),
                                                 ctx->counts_stage, 0, PTW_INVALID_VERTEX, 0, s));
    // a copy the size of a counts output takes the path walks take (small
    // copies and large ones use different engines)
    HIP_TRY(ctx, hipMemcpyAsync(ctx->pinned, ctx->scratch, 8 * sizeof(uint64_t), hipMemcpyDeviceToHost, s));
    HIP_TRY(ctx, hipMemcpyAsync(ctx->pinned + 8, ctx->counts_stage, (size_t)ctx->dev.V * 8, hipMemcpyDeviceToHost, s));
    HIP_TRY(ctx, hipMemcpyAsync(ctx->counts_stage, ctx->pinned + 8, (size_t)ctx->dev.V * 8, hipMemcpyHostToDevice, s));
    HIP_TRY(ctx, hipStreamSynchronize(s));
    // One real walk of one walker through ptw_walk: the first walk of a
    // process otherwise pays about 0.3 ms of one-time host-side set-up that
    // the launches above do not take (measured: 0.54 ms for the first
    // 655,360-walker PageRank on R-MAT scale 16, 0.26 ms for the next ones).
    if (ctx->dev.V) {
        ptw_walk_desc d;
        std::memset(&d, 0, sizeof(d));
        d.app = PTW_APP_PAGERANK;
        d.semantics = PTW_SEM_REFERENCE;
        d.start_mode = PTW_START_UNIFORM;
        d.dangling_policy = PTW_DANGLING_STOP;
        d.tp = 1.0;
        d.seed = 1;
        d.walkers = 1;
        d.walk_len = 1;
        std::vector<uint64_t> counts(ctx->dev.V);
        ptw_walk_out out;
        std::memset(&out, 0, sizeof(out));
        out.counts = counts.data();
        if (int rc = ptw_walk(ctx, &d, &out)) return rc;
    }
    return PTW_OK;
}

// The node2vec tables, built on first use: whether the graph is simple and
// symmetric (checked on the host over rows sorted by device id) and the
// device edge set that answers "is x a neighbour of prev".
int ensure_node2vec_tables(ptw_ctx *ctx)
{
    ptw::DeviceGraph &d = ctx->dev;
    if (d.eset) return PTW_OK;
    const HostCsr &h = ctx->host;
    const uint32_t V = d.V;
    std::vector<uint32_t> scol(h.E());
    const bool rel = d.relabelled;
    ptw::parallel_for(V, [&](uint64_t lo, uint64_t hi) {
        for (uint64_t u = lo; u < hi; ++u) {
            const uint32_t g = rel ? ctx->h_inv[u] : (uint32_t)u;
            const uint32_t b = h.row_ptr[g], e = h.row_ptr[g + 1];
            const uint32_t out = rel ? ctx->h_drow[u] : b;
            for (uint32_t k = b; k < e; ++k) scol[out + (k - b)] = rel ? ctx->h_perm[h.col[k]] : h.col[k];
            std::sort(scol.begin() + out, scol.begin() + out + (e - b));
        }
    });
    // simple and symmetric: no sorted row repeats a neighbour, and every
    // edge u -> x finds u in x's sorted row
    const uint32_t *ro = rel ? ctx->h_drow.data() : h.row_ptr.data();
    std::atomic<bool> ok{true};
    ptw::parallel_for(V, [&](uint64_t lo, uint64_t hi) {
        for (uint64_t u = lo; u < hi && ok.load(std::memory_order_relaxed); ++u) {
            const uint32_t *a = scol.data() + ro[u];
            const uint32_t n = ro[u + 1] - ro[u];
            for (uint32_t k = 0; k < n; ++k) {
                const uint32_t *b = scol.data() + ro[a[k]];
                if ((k && a[k] == a[k - 1]) || !std::binary_search(b, b + (ro[a[k] + 1] - ro[a[k]]), (uint32_t)u)) {
                    ok.store(false, std::memory_order_relaxed);
                    break;
                }
            }
        }
    });
    d.simple_symmetric = ok.load();
    std::vector<uint32_t>().swap(scol);
    // open addressing at most half full: the smallest power of two >= 2E slots
    const uint64_t E = h.E();
    uint64_t slots = 2;
    while (slots < 2 * E) slots <<= 1;
    struct Tmp {
        void *p = nullptr;
        ~Tmp() { if (p) (void)hipFree(p); }
    } t;
    HIP_TRY(ctx, hipMalloc(&t.p, slots * sizeof(unsigned long long)));
    unsigned long long *table = (unsigned long long *)t.p;
    HIP_TRY(ctx, ptw::launch_fill_u64(table, slots, ptw::kEdgeSetEmpty, ctx->stream));
    HIP_TRY(ctx, ptw::launch_edge_set_build(d.row_ptr, d.col, V, E, table, slots - 1, ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    t.p = nullptr;
    d.eset = table;
    d.eset_mask = slots - 1;
    d.bytes += slots * sizeof(unsigned long long);
    return PTW_OK;
}

// With `built`, the device graph was made on the device (ingest) and is taken
// over as it is, with its relabel; otherwise it is relabelled here and uploaded.
int install_graph(ptw_ctx *ctx, HostCsr &&g, uint32_t flags, ptw::DeviceBuiltGraph *built = nullptr)
{
    free_graph(ctx);
    ctx->host = std::move(g);
    HostCsr &h = ctx->host;
    const uint32_t V = h.V;
    // degree statistics
    uint32_t maxd = 0, maxv = 0, dang = 0;
    for (uint32_t v = 0; v < V; ++v) {
        uint32_t d = h.row_ptr[v + 1] - h.row_ptr[v];
        if (d > maxd) {
            maxd = d;
            maxv = v;
        }
        dang += d == 0;
    }
    ctx->max_degree = maxd;
    ctx->max_degree_vertex = maxv;
    ctx->num_dangling = dang;

    auto &d = ctx->dev;
    d.V = V;
    d.E = h.E();
    if (ctx->device < 0) {  // host-only context: no device copy
        ctx->has_graph = true;
        return PTW_OK;
    }
    int rc;
    if (built) {
        d.row_ptr = built->row_ptr;
        d.col = built->col;
        d.perm = built->perm;
        d.inv = built->inv;
        d.relabelled = built->relabelled;
        d.hubs = built->relabelled ? V : 0;
        ctx->h_perm = std::move(built->h_perm);
        ctx->h_inv = std::move(built->h_inv);
        ctx->h_drow = std::move(built->h_drow);
        *built = ptw::DeviceBuiltGraph{};
    } else if (flags & PTW_GRAPH_NO_RELABEL) {
        if ((rc = upload(ctx, &d.row_ptr, h.row_ptr.data(), h.row_ptr.size()))) return rc;
        if ((rc = upload(ctx, &d.col, h.col.data(), h.col.size()))) return rc;
        d.relabelled = false;
        d.hubs = 0;
    } else {
        std::vector<uint32_t> col;
        ptw::relabel_by_degree(h, ctx->h_perm, ctx->h_inv, ctx->h_drow, col);
        if ((rc = upload(ctx, &d.row_ptr, ctx->h_drow.data(), ctx->h_drow.size()))) return rc;
        if ((rc = upload(ctx, &d.col, col.data(), col.size()))) return rc;
        if ((rc = upload(ctx, &d.perm, ctx->h_perm.data(), ctx->h_perm.size()))) return rc;
        if ((rc = upload(ctx, &d.inv, ctx->h_inv.data(), ctx->h_inv.size()))) return rc;
        d.relabelled = true;
        d.hubs = V;
    }
    HIP_TRY(ctx, hipMalloc((void **)&ctx->counts_dev, (size_t)(V ? V : 1) * sizeof(unsigned long long)));
    HIP_TRY(ctx, hipMalloc((void **)&ctx->counts_stage, (size_t)(V ? V : 1) * sizeof(uint64_t)));
    // count outputs cross PCIe through page-locked memory: a pageable copy
    // pays a staging setup on first use and runs at a fraction of the link
    HIP_TRY(ctx, hipHostMalloc((void **)&ctx->pinned, (8 + (size_t)V) * sizeof(uint64_t), hipHostMallocDefault));
    d.bytes = ((uint64_t)V + 1) * 4 + d.E * 4 + (d.relabelled ? 8ull * V : 0) + 8ull * V;
    ctx->has_graph = true;
    return warm_up(ctx);
}

double now_ms()
{
    using namespace std::chrono;
    return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

// Descriptor and output checks shared by ptw_walk and ptw_walk_static.
int check_walk(ptw_ctx *ctx, const ptw_walk_desc &d, const ptw_walk_out &out)
{
    if (!ctx->has_graph) return fail(ctx, PTW_ERR_NO_GRAPH, "no graph loaded");
    if (ctx->device < 0) return fail(ctx, PTW_ERR_UNSUPPORTED, "host-only context: walks need a HIP device");
    const uint32_t V = ctx->dev.V;
    if (d.app < PTW_APP_PAGERANK || d.app > PTW_APP_NODE2VEC) return fail(ctx, PTW_ERR_INVALID_ARG, "unknown app");
    const bool path_app = is_path_app(d.app), count_app = is_count_app(d.app);
    if (d.semantics != PTW_SEM_REFERENCE && d.semantics != PTW_SEM_TEXTBOOK)
        return fail(ctx, PTW_ERR_INVALID_ARG, "unknown semantics");
    if (d.start_mode < PTW_START_UNIFORM || d.start_mode > PTW_START_ROUNDROBIN)
        return fail(ctx, PTW_ERR_INVALID_ARG, "unknown start mode");
    if (d.dangling_policy != PTW_DANGLING_ERROR && d.dangling_policy != PTW_DANGLING_STOP)
        return fail(ctx, PTW_ERR_INVALID_ARG, "unknown dangling policy");
    if (!path_app && !(d.tp > 0.0 && d.tp <= 1.0))
        return fail(ctx, PTW_ERR_INVALID_ARG, "termination probability must be in (0, 1]");
    if (path_app && d.walk_len == 0) return fail(ctx, PTW_ERR_INVALID_ARG, "walk length must be >= 1");
    if (d.app == PTW_APP_NODE2VEC) {
        // attempts are counted in bits 20.. of the step word
        if (d.walk_len > (1u << 20)) return fail(ctx, PTW_ERR_INVALID_ARG, "node2vec walk length must be <= 2^20");
        const double lo = 1.0 / 65536.0, hi = 65536.0;
        if (!(d.p >= lo && d.p <= hi) || !(d.q >= lo && d.q <= hi))
            return fail(ctx, PTW_ERR_INVALID_ARG, "node2vec p and q must be in [2^-16, 2^16]");
        if (d.flags & PTW_FLAG_WEIGHTED)
            return fail(ctx, PTW_ERR_UNSUPPORTED, "node2vec on weighted graphs is not supported");
    }
    if (d.start_mode == PTW_START_SOURCE && d.source >= V)
        return fail(ctx, PTW_ERR_INVALID_ARG, "source vertex " + std::to_string(d.source) + " out of range");
    if (count_app && !out.counts) return fail(ctx, PTW_ERR_INVALID_ARG, "counts output required");
    if (d.app == PTW_APP_RWR && !out.ends) return fail(ctx, PTW_ERR_INVALID_ARG, "ends output required");
    if (d.app != PTW_APP_RWR && out.starts)
        return fail(ctx, PTW_ERR_INVALID_ARG, "starts output is for RWR walks only");
    if (path_app && !out.paths) return fail(ctx, PTW_ERR_INVALID_ARG, "paths output required");
    if (path_app && d.walkers > UINT64_MAX / 4 / d.walk_len)
        return fail(ctx, PTW_ERR_INVALID_ARG, "path buffer size overflows");
    if (d.wid_base + d.walkers < d.wid_base) return fail(ctx, PTW_ERR_INVALID_ARG, "walker id range overflows");
    if ((d.flags & PTW_FLAG_WEIGHTED) && !ctx->dev.wedge)
        return fail(ctx, PTW_ERR_INVALID_ARG, "weighted walk on a graph without weights");
    if ((d.flags & PTW_FLAG_OUT_DEVICE) &&
        (((uintptr_t)out.paths & 15) || ((uintptr_t)out.ends & 3) || ((uintptr_t)out.starts & 3) ||
         ((uintptr_t)out.counts & 7)))
        return fail(ctx, PTW_ERR_INVALID_ARG,
                    "device outputs must be aligned: paths 16 B, ends and starts 4 B, counts 8 B");
    return PTW_OK;
}

// Launch parameters from the descriptor: graph, process, node2vec weights.
// Outputs, counters and scratch are the caller's to set.
int make_walk_launch(ptw_ctx *ctx, const ptw_walk_desc &d, ptw::WalkLaunch &p)
{
    p = ptw::WalkLaunch{};
    p.row_ptr = ctx->dev.row_ptr;
    p.col = ctx->dev.col;
    p.perm = ctx->dev.perm;
    p.inv = ctx->dev.inv;
    p.wedge = (d.flags & PTW_FLAG_WEIGHTED) ? ctx->dev.wedge : nullptr;
    p.V = ctx->dev.V;
    p.hubs = ctx->dev.hubs;
    p.app = d.app;
    p.semantics = d.semantics;
    p.start_mode = d.start_mode;
    p.source_dev = d.start_mode == PTW_START_SOURCE && ctx->dev.relabelled ? ctx->h_perm[d.source] : d.source;
    p.walk_len = d.walk_len;
    p.thr = ptw::continue_threshold(d.tp);
    p.seed_lo = (uint32_t)d.seed;
    p.seed_hi = (uint32_t)(d.seed >> 32);
    p.stream = d.stream;
    p.wid_base = d.wid_base;
    p.walkers = d.walkers;
    if (d.app == PTW_APP_NODE2VEC) {
        if (int rc = ensure_node2vec_tables(ctx)) return rc;
        p.eset = ctx->dev.eset;
        p.eset_mask = ctx->dev.eset_mask;
        // Integer weights floor(2^32 w / M).  Unfolded, M = max(1/p, 1, 1/q)
        // and every weight is an acceptance threshold in [0, 2^32].  Folded
        // (1/p above both other weights on a simple symmetric graph), M =
        // max(1, 1/q) and the return weight's excess over 2^32 is split over
        // ceil(excess / 2^32) extra proposal slots.
        const double wr = 1.0 / d.p, wo = 1.0 / d.q;
        const double Mf = std::max(1.0, wo);
        const uint64_t t_ret = (uint64_t)std::floor(wr / Mf * 4294967296.0);  // <= 2^48
        const bool fold = ctx->dev.simple_symmetric && wr > Mf && t_ret > 4294967296ull &&
                          ctx->max_degree < (1u << 31);
        const double M = fold ? Mf : std::max(wr, Mf);
        auto thr = [M](double w) {
            const double t = std::floor(w / M * 4294967296.0);
            return t >= 4294967296.0 ? 4294967296ull : (uint64_t)t;
        };
        p.w_in = thr(1.0);
        p.w_out = thr(wo);
        if (fold) {
            const uint64_t excess = t_ret - 4294967296ull;
            const uint64_t slots = (excess + 4294967295ull) >> 32;  // in [1, 2^16]
            p.n2v_fold = (uint32_t)slots;
            p.w_fold = excess / slots;  // <= 2^32
            p.w_ret = 4294967296ull + slots * p.w_fold;
        } else {
            p.n2v_fold = 0;
            p.w_fold = 0;
            p.w_ret = thr(wr);
        }
    }
    return PTW_OK;
}

// The context's launch shape with cache_slots = PTW_CACHE_AUTO resolved for
// this walk.  The visit cache pays only for the hot neighbourhood of a single
// source.  Uniform starts spread their visits, and a source with more
// neighbours than four times the cache spreads its first-step visits too thin
// (about two per workgroup per neighbour at 1e8 walkers from a 1.6e5-degree
// hub), so there the cache's LDS goes to hub counters (4 per slot).  A walk
// that counts through the visit log gives the cache up too (launch_walk).
// An explicit cache_slots is used as given.
ptw::LaunchShape walk_shape(const ptw_ctx *ctx, const ptw_walk_desc &d)
{
    ptw::LaunchShape shape = ctx->shape;
    if (shape.cache_slots < 0) {
        const bool keep_cache = d.start_mode == PTW_START_SOURCE &&
                                (uint64_t)ctx->host.row_ptr[d.source + 1] - ctx->host.row_ptr[d.source] <=
                                    4ull * ptw::kAutoCacheSlots;
        if (is_count_app(d.app) && !keep_cache) {
            shape.lds_hubs = std::min(ptw::kMaxLdsHubs, shape.lds_hubs + 4 * ptw::kAutoCacheSlots);
            shape.cache_slots = 0;
        } else {
            shape.cache_slots = ptw::kAutoCacheSlots;
            shape.cache_auto = true;
        }
    }
    return shape;
}

void clear_walk_out(ptw_walk_out *out)
{
    out->steps = out->visits = out->dangling = 0;
    out->kernel_ms = out->total_ms = 0.0;
    out->launch_block = out->launch_blocks = out->launch_chunk = 0;
    out->launch_lds_hubs = out->launch_cache_slots = 0;
    out->launch_visit_log = 0;
}

// Totals from the scratch words (copied to host: [1] steps, [2] dangling),
// the kernel time and the launch shape, then the dangling-walker verdict.
int finish_walk_out(ptw_ctx *ctx, const ptw_walk_desc &d, const uint64_t *sc, const ptw::EffectiveShape &eff,
                    double t0, ptw_walk_out *out, uint64_t vlog_cap = 0)
{
    float ms = 0.f;
    HIP_TRY(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    out->kernel_ms = ms;
    out->steps = sc[1];
    out->dangling = sc[2];
    out->visits = is_count_app(d.app) ? d.walkers + sc[1] : 0;
    out->total_ms = now_ms() - t0;
    out->launch_block = (uint32_t)eff.block;
    out->launch_blocks = (uint32_t)eff.blocks;
    out->launch_chunk = eff.chunk;
    out->launch_lds_hubs = eff.hubs;
    out->launch_cache_slots = eff.slots;
    out->launch_visit_log = eff.vlog ? (uint32_t)std::min<uint64_t>(sc[3], vlog_cap) : 0;
    if (sc[2] && d.dangling_policy == PTW_DANGLING_ERROR)
        return fail(ctx, PTW_ERR_DANGLING,
                    std::to_string(sc[2]) + " walkers reached a vertex without out-edges with steps left");
    return PTW_OK;
}

}  // namespace

extern "C" {

int ptw_version(void) { return PTW_VERSION_MAJOR * 1000 + PTW_VERSION_MINOR; }

const char *ptw_last_error(const ptw_ctx *ctx)
{
    if (!ctx) return "null context";
    return ctx->err.c_str();
}

int ptw_create(ptw_ctx **out, int device)
{
    if (!out) return PTW_ERR_INVALID_ARG;
    *out = nullptr;
    ptw_ctx *ctx = new (std::nothrow) ptw_ctx;
    if (!ctx) return PTW_ERR_OOM;
    if (device == PTW_DEVICE_HOST_ONLY) {
        ctx->device = -1;  // graph ingest only; walks report PTW_ERR_UNSUPPORTED
        *out = ctx;
        return PTW_OK;
    }
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n == 0) {
        delete ctx;
        return PTW_ERR_HIP;  // no HIP device: the engine has no CPU path
    }
    if (device < 0 || device >= n) {
        delete ctx;
        return PTW_ERR_INVALID_ARG;
    }
    ctx->device = device;
    if (hipSetDevice(device) != hipSuccess ||
        hipDeviceGetAttribute(&ctx->num_cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess ||
        hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&ctx->ev0) != hipSuccess || hipEventCreate(&ctx->ev1) != hipSuccess ||
        hipMalloc((void **)&ctx->scratch, 8 * sizeof(unsigned long long)) != hipSuccess ||
        ptw::static_scratch_address(device, &ctx->static_scratch) != hipSuccess ||
        // first launch loads the device code, so no walk pays for it
        ptw::launch_fill_u64(ctx->scratch, 8, 0ull, ctx->stream) != hipSuccess ||
        hipStreamSynchronize(ctx->stream) != hipSuccess) {
        ptw_destroy(ctx);
        return PTW_ERR_HIP;
    }
    *out = ctx;
    return PTW_OK;
}

void ptw_destroy(ptw_ctx *ctx)
{
    if (!ctx) return;
    if (ctx->device < 0) {
        ctx->host = HostCsr{};
        delete ctx;
        return;
    }
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    ptw::destroy_serve_lanes(ctx);
    free_graph(ctx);
    if (ctx->scratch) (void)hipFree(ctx->scratch);
    if (ctx->ev0) (void)hipEventDestroy(ctx->ev0);
    if (ctx->ev1) (void)hipEventDestroy(ctx->ev1);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

int ptw_graph_load(ptw_ctx *ctx, const char *path, int format, int as_undirected, uint32_t flags)
{
    if (!ctx) return PTW_ERR_INVALID_ARG;
    try {
        if (ctx->device >= 0) (void)hipSetDevice(ctx->device);
        if ((flags & PTW_GRAPH_INGEST_HOST) && (flags & PTW_GRAPH_INGEST_DEVICE))
            return fail(ctx, PTW_ERR_INVALID_ARG, "PTW_GRAPH_INGEST_HOST and PTW_GRAPH_INGEST_DEVICE exclude each other");
        bool on_device = (flags & PTW_GRAPH_INGEST_DEVICE) != 0;
        if (on_device && ctx->device < 0)
            return fail(ctx, PTW_ERR_UNSUPPORTED, "host-only context: device ingest needs a HIP device");
        if (!on_device && !(flags & PTW_GRAPH_INGEST_HOST) && ctx->device >= 0 && path) {
            struct stat st;
            on_device = stat(path, &st) == 0 && (uint64_t)st.st_size >= (16ull << 20);
        }
        HostCsr g;
        if (on_device) {
            ptw::DeviceBuiltGraph built;
            int rc = ptw::device_load_edge_list(ctx->stream, path, format, as_undirected != 0,
                                                !(flags & PTW_GRAPH_NO_RELABEL), g, built, ctx->err);
            if (rc) return rc;
            if (g.V == 0)
                return fail(ctx, PTW_ERR_FORMAT, std::string("graph file '") + path + "' holds no edges");
            return install_graph(ctx, std::move(g), flags, &built);
        }
        int rc = ptw::load_edge_list(path, format, as_undirected != 0, g, ctx->err);
        if (rc) return rc;
        if (g.V == 0) return fail(ctx, PTW_ERR_FORMAT, std::string("graph file '") + (path ? path : "") + "' holds no edges");
        return install_graph(ctx, std::move(g), flags);
    } catch (const std::bad_alloc &) {
        return fail(ctx, PTW_ERR_OOM, "host allocation failed while loading the graph");
    } catch (const std::exception &ex) {
        return fail(ctx, PTW_ERR_IO, ex.what());
    }
}

int ptw_graph_from_csr(ptw_ctx *ctx, const uint32_t *row_ptr, const uint32_t *col, uint32_t num_vertices,
                       uint64_t num_edges, uint32_t flags)
{
    if (!ctx) return PTW_ERR_INVALID_ARG;
    try {
        if (ctx->device >= 0) (void)hipSetDevice(ctx->device);
        int rc = ptw::validate_csr(row_ptr, col, num_vertices, num_edges, ctx->err);
        if (rc) return rc;
        HostCsr g;
        g.V = num_vertices;
        g.row_ptr.assign(row_ptr, row_ptr + num_vertices + 1);
        g.col.assign(col, col + num_edges);
        return install_graph(ctx, std::move(g), flags);
    } catch (const std::bad_alloc &) {
        return fail(ctx, PTW_ERR_OOM, "host allocation failed while copying the CSR");
    }
}

int ptw_graph_rmat(ptw_ctx *ctx, uint32_t scale, uint32_t edge_factor, double a, double b, double c,
                   uint64_t seed, uint32_t flags)
{
    if (!ctx) return PTW_ERR_INVALID_ARG;
    try {
        if (ctx->device >= 0) (void)hipSetDevice(ctx->device);
        HostCsr g;
        int rc = ptw::build_rmat(scale, edge_factor, a, b, c, seed, (flags & PTW_RMAT_KEEP_ISOLATED) != 0, g,
                                 ctx->err);
        if (rc) return rc;
        if (g.V == 0) return fail(ctx, PTW_ERR_INVALID_ARG, "R-MAT produced no edges");
        return install_graph(ctx, std::move(g), flags);
    } catch (const std::bad_alloc &) {
        return fail(ctx, PTW_ERR_OOM, "host allocation failed while generating R-MAT");
    }
}

int ptw_graph_set_weights(ptw_ctx *ctx, const float *weights, uint64_t num_edges)
{
    if (!ctx || (num_edges && !weights)) return fail(ctx, PTW_ERR_INVALID_ARG, "null argument");
    if (!ctx->has_graph) return fail(ctx, PTW_ERR_NO_GRAPH, "no graph loaded");
    if (num_edges != ctx->host.E()) return fail(ctx, PTW_ERR_INVALID_ARG, "weight count differs from the edge count");
    try {
        if (ctx->device >= 0) HIP_TRY(ctx, hipSetDevice(ctx->device));
        ctx->weights.assign(weights, weights + num_edges);
        int rc = install_weights(ctx);
        if (rc) ctx->weights.clear();
        return rc;
    } catch (const std::bad_alloc &) {
        return fail(ctx, PTW_ERR_OOM, "host allocation failed while building alias tables");
    }
}

int ptw_graph_random_weights(ptw_ctx *ctx, uint64_t seed)
{
    if (!ctx) return PTW_ERR_INVALID_ARG;
    if (!ctx->has_graph) return fail(ctx, PTW_ERR_NO_GRAPH, "no graph loaded");
    try {
        if (ctx->device >= 0) HIP_TRY(ctx, hipSetDevice(ctx->device));
        ptw::random_weights(ctx->host.E(), seed, ctx->weights);
        return install_weights(ctx);
    } catch (const std::bad_alloc &) {
        return fail(ctx, PTW_ERR_OOM, "host allocation failed while building alias tables");
    }
}

int ptw_graph_export_weights(ptw_ctx *ctx, float *weights, uint32_t *alias_thr, uint32_t *alias_idx)
{
    if (!ctx) return PTW_ERR_INVALID_ARG;
    if (!ctx->has_graph) return fail(ctx, PTW_ERR_NO_GRAPH, "no graph loaded");
    if (ctx->weights.empty() && ctx->host.E()) return fail(ctx, PTW_ERR_NO_GRAPH, "graph has no weights");
    const size_t E = ctx->host.E();
    if (weights) memcpy(weights, ctx->weights.data(), E * 4);
    if (alias_thr) memcpy(alias_thr, ctx->alias_thr.data(), E * 4);
    if (alias_idx) memcpy(alias_idx, ctx->alias_idx.data(), E * 4);
    return PTW_OK;
}

int ptw_graph_info_get(ptw_ctx *ctx, ptw_graph_info *info)
{
    if (!ctx || !info) return fail(ctx, PTW_ERR_INVALID_ARG, "null argument");
    if (!ctx->has_graph) return fail(ctx, PTW_ERR_NO_GRAPH, "no graph loaded");
    info->num_vertices = ctx->dev.V;
    info->num_edges = ctx->dev.E;
    info->max_degree = ctx->max_degree;
    info->max_degree_vertex = ctx->max_degree_vertex;
    info->num_dangling = ctx->num_dangling;
    info->hub_vertices = ctx->dev.hubs < (uint32_t)ctx->shape.lds_hubs ? ctx->dev.hubs : (uint32_t)ctx->shape.lds_hubs;
    info->relabelled = ctx->dev.relabelled ? 1 : 0;
    info->device_bytes = ctx->dev.bytes;
    info->weighted = ctx->weights.empty() ? 0 : 1;
    return PTW_OK;
}

int ptw_graph_export_names(ptw_ctx *ctx, uint32_t *names)
{
    if (!ctx) return PTW_ERR_INVALID_ARG;
    if (!ctx->has_graph) return fail(ctx, PTW_ERR_NO_GRAPH, "no graph loaded");
    if (!names) return fail(ctx, PTW_ERR_INVALID_ARG, "null output buffer");
    const HostCsr &h = ctx->host;
    if (h.names.empty())
        for (uint32_t v = 0; v < h.V; ++v) names[v] = v;
    else
        memcpy(names, h.names.data(), (size_t)h.V * 4);
    return PTW_OK;
}

int ptw_graph_export_csr(ptw_ctx *ctx, uint32_t *row_ptr, uint32_t *col)
{
    if (!ctx) return PTW_ERR_INVALID_ARG;
    if (!ctx->has_graph) return fail(ctx, PTW_ERR_NO_GRAPH, "no graph loaded");
    if (!row_ptr || (ctx->host.E() && !col)) return fail(ctx, PTW_ERR_INVALID_ARG, "null output buffer");
    memcpy(row_ptr, ctx->host.row_ptr.data(), ctx->host.row_ptr.size() * 4);
    if (ctx->host.E()) memcpy(col, ctx->host.col.data(), ctx->host.col.size() * 4);
    return PTW_OK;
}

int ptw_graph_write(ptw_ctx *ctx, const char *path, int format)
{
    if (!ctx || !path) return fail(ctx, PTW_ERR_INVALID_ARG, "null argument");
    if (!ctx->has_graph) return fail(ctx, PTW_ERR_NO_GRAPH, "no graph loaded");
    if (format != PTW_FORMAT_TEXT && format != PTW_FORMAT_BINARY)
        return fail(ctx, PTW_ERR_INVALID_ARG, "unknown graph format");
    return ptw::write_edge_list(ctx->host, path, format, ctx->err);
}

int ptw_walk(ptw_ctx *ctx, const ptw_walk_desc *desc, ptw_walk_out *out)
{
    if (!ctx) return PTW_ERR_INVALID_ARG;
    if (!desc || !out) return fail(ctx, PTW_ERR_INVALID_ARG, "null descriptor or output");
    const double t0 = now_ms();
    const ptw_walk_desc &d = *desc;
    if (int rc = check_walk(ctx, d, *out)) return rc;
    const uint32_t V = ctx->dev.V;
    const bool count_app = is_count_app(d.app), path_app = is_path_app(d.app);
    const bool on_dev = (d.flags & PTW_FLAG_OUT_DEVICE) != 0;
    const bool accumulate = (d.flags & PTW_FLAG_ACCUMULATE) != 0;
    clear_walk_out(out);
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    hipStream_t s = ctx->stream;

    // device-side output buffers
    uint32_t *ends_dev = nullptr, *paths_dev = nullptr, *starts_dev = nullptr;
    uint64_t *counts_out_dev = nullptr;
    struct Tmp {
        void *p = nullptr;
        ~Tmp() { if (p) (void)hipFree(p); }
    } t_ends, t_paths, t_starts;
    const uint64_t W = d.walkers;
    if (d.app == PTW_APP_RWR) {
        if (on_dev) {
            ends_dev = out->ends;
            starts_dev = out->starts;
        } else if (W) {
            HIP_TRY(ctx, hipMalloc(&t_ends.p, W * 4));
            ends_dev = (uint32_t *)t_ends.p;
            if (out->starts) {
                HIP_TRY(ctx, hipMalloc(&t_starts.p, W * 4));
                starts_dev = (uint32_t *)t_starts.p;
            }
        }
    }
    if (path_app) {
        if (on_dev) {
            paths_dev = out->paths;
        } else if (W) {
            HIP_TRY(ctx, hipMalloc(&t_paths.p, W * d.walk_len * 4));
            paths_dev = (uint32_t *)t_paths.p;
        }
    }
    if (count_app) {
        if (on_dev) {
            counts_out_dev = out->counts;
        } else {
            counts_out_dev = ctx->counts_stage;
            if (accumulate) {
                std::memcpy(ctx->pinned + 8, out->counts, (size_t)V * 8);
                HIP_TRY(ctx, hipMemcpyAsync(counts_out_dev, ctx->pinned + 8, (size_t)V * 8, hipMemcpyHostToDevice, s));
            }
        }
    }

    ptw::WalkLaunch p;
    if (int rc = make_walk_launch(ctx, d, p)) return rc;
    pick_edge_records(ctx, d, true, p);
    p.counts_dev = ctx->counts_dev;
    p.ends = ends_dev;
    p.starts = starts_dev;
    p.paths = paths_dev;
    p.scratch = ctx->scratch;

    HIP_TRY(ctx, ptw::launch_fill_u64(ctx->scratch, 8, 0ull, s));
    if (count_app) HIP_TRY(ctx, ptw::launch_fill_u64(ctx->counts_dev, V, 0ull, s));
    const ptw::LaunchShape shape = walk_shape(ctx, d);
    // With an automatic launch, a walk under 8 walkers per lane of one
    // workgroup per CU counts with device atomics (launch_walk's small-walk
    // rule; a count launch holds one workgroup per CU), so it takes no log
    // and allocates none.
    const bool small_walk =
        shape.blocks_per_cu <= 0 && W < 8ull * (uint64_t)ctx->num_cus * (uint64_t)shape.block;
    const uint64_t log_cap =
        count_app && W && shape.visit_log && !small_walk
            ? ensure_vlog(ctx, W, d.tp, shape.visit_log > 1 ? (uint64_t)shape.visit_log : 0)
            : 0;
    if (log_cap) {
        p.vlog = ctx->vlog.log;
        p.vlog_cap = log_cap;
        p.vlog_hist = ctx->vlog.hist();
        p.vlog_shift = ptw::vlog_shift_for(V);
        p.vlog_buckets = ptw::vlog_buckets_for(V, p.vlog_shift);
        HIP_TRY(ctx, ptw::launch_vlog_begin(ctx->vlog, p.vlog_buckets, s));
    }
    HIP_TRY(ctx, hipEventRecord(ctx->ev0, s));
    ptw::EffectiveShape eff;
    if (W) HIP_TRY(ctx, ptw::launch_walk(p, shape, ctx->num_cus, s, &eff));
    // the log's counting belongs to the walk's device time
    if (eff.vlog)
        HIP_TRY(ctx, ptw::launch_vlog_count(ctx->vlog, ctx->scratch + 3, log_cap, p.vlog_shift, p.vlog_buckets,
                                            ctx->counts_dev, ctx->num_cus, s));
    HIP_TRY(ctx, hipEventRecord(ctx->ev1, s));
    if (count_app) {
        const bool from_source = d.start_mode == PTW_START_SOURCE;
        HIP_TRY(ctx, ptw::launch_counts_to_graph_ids(ctx->counts_dev, ctx->dev.perm, V, counts_out_dev, accumulate,
                                                     from_source ? d.source : PTW_INVALID_VERTEX,
                                                     from_source ? W : 0, s));
    }
    HIP_TRY(ctx, hipMemcpyAsync(ctx->pinned, ctx->scratch, 8 * sizeof(uint64_t), hipMemcpyDeviceToHost, s));
    if (!on_dev) {
        if (count_app)
            HIP_TRY(ctx, hipMemcpyAsync(ctx->pinned + 8, counts_out_dev, (size_t)V * 8, hipMemcpyDeviceToHost, s));
        if (ends_dev && W) HIP_TRY(ctx, hipMemcpyAsync(out->ends, ends_dev, W * 4, hipMemcpyDeviceToHost, s));
        if (starts_dev && W) HIP_TRY(ctx, hipMemcpyAsync(out->starts, starts_dev, W * 4, hipMemcpyDeviceToHost, s));
        if (paths_dev && W)
            HIP_TRY(ctx, hipMemcpyAsync(out->paths, paths_dev, W * d.walk_len * 4, hipMemcpyDeviceToHost, s));
    }
    HIP_TRY(ctx, hipStreamSynchronize(s));
    if (!on_dev && count_app) std::memcpy(out->counts, ctx->pinned + 8, (size_t)V * 8);
    return finish_walk_out(ctx, d, ctx->pinned, eff, t0, out, log_cap);
}

int ptw_walk_static(ptw_ctx *ctx, const ptw_walk_desc *desc, ptw_walk_out *out)
{
    if (!ctx) return PTW_ERR_INVALID_ARG;
    if (!desc || !out) return fail(ctx, PTW_ERR_INVALID_ARG, "null descriptor or output");
    const double t0 = now_ms();
    const ptw_walk_desc &d = *desc;
    if (int rc = check_walk(ctx, d, *out)) return rc;
    const uint32_t V = ctx->dev.V;
    const bool count_app = is_count_app(d.app);
    if (!(d.flags & PTW_FLAG_OUT_DEVICE))
        return fail(ctx, PTW_ERR_INVALID_ARG, "ptw_walk_static writes device outputs only: set PTW_FLAG_OUT_DEVICE");
    if (d.app == PTW_APP_NODE2VEC)
        return fail(ctx, PTW_ERR_UNSUPPORTED,
                    "ptw_walk_static does not run node2vec (its edge set is allocated on first use): use ptw_walk");
    if (count_app && V > ptw::kStaticMaxVertices)
        return fail(ctx, PTW_ERR_INVALID_ARG,
                    "ptw_walk_static counts at most " + std::to_string(ptw::kStaticMaxVertices) +
                        " vertices in its static scratch; the graph has " + std::to_string(V) + ": use ptw_walk");
    if (!ctx->static_scratch) return fail(ctx, PTW_ERR_HIP, "no static scratch address for this device");
    clear_walk_out(out);
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    hipStream_t s = ctx->stream;
    // [0, 8): queue and totals; [8, 8 + V): device-id counters.  The array
    // keeps the last call's contents, so both are cleared first.
    unsigned long long *const words = ctx->static_scratch;
    unsigned long long *const counters = words + 8;

    ptw::WalkLaunch p;
    if (int rc = make_walk_launch(ctx, d, p)) return rc;
    pick_edge_records(ctx, d, false, p);  // only records built by earlier walks: no allocation here
    p.counts_dev = counters;
    p.ends = d.app == PTW_APP_RWR ? out->ends : nullptr;
    p.starts = d.app == PTW_APP_RWR ? out->starts : nullptr;
    p.paths = is_path_app(d.app) ? out->paths : nullptr;
    p.scratch = words;

    HIP_TRY(ctx, ptw::launch_fill_u64(words, 8, 0ull, s));
    if (count_app) HIP_TRY(ctx, ptw::launch_fill_u64(counters, V, 0ull, s));
    HIP_TRY(ctx, hipEventRecord(ctx->ev0, s));
    const ptw::LaunchShape shape = walk_shape(ctx, d);
    ptw::EffectiveShape eff;
    if (d.walkers) HIP_TRY(ctx, ptw::launch_walk(p, shape, ctx->num_cus, s, &eff));
    HIP_TRY(ctx, hipEventRecord(ctx->ev1, s));
    if (count_app) {
        const bool from_source = d.start_mode == PTW_START_SOURCE;
        HIP_TRY(ctx, ptw::launch_counts_to_graph_ids(counters, ctx->dev.perm, V, out->counts,
                                                     (d.flags & PTW_FLAG_ACCUMULATE) != 0,
                                                     from_source ? d.source : PTW_INVALID_VERTEX,
                                                     from_source ? d.walkers : 0, s));
    }
    // the totals travel through the page-locked words the graph install made
    HIP_TRY(ctx, hipMemcpyAsync(ctx->pinned, words, 8 * sizeof(uint64_t), hipMemcpyDeviceToHost, s));
    HIP_TRY(ctx, hipStreamSynchronize(s));
    return finish_walk_out(ctx, d, ctx->pinned, eff, t0, out);
}

int ptw_tuning_get(ptw_ctx *ctx, ptw_tuning *t)
{
    if (!ctx || !t) return fail(ctx, PTW_ERR_INVALID_ARG, "null argument");
    t->block = ctx->shape.block;
    t->blocks_per_cu = ctx->shape.blocks_per_cu;
    t->chunk = ctx->shape.chunk;
    t->lds_hubs = ctx->shape.lds_hubs;
    t->cache_slots = ctx->shape.cache_slots;
    t->batch_chunk = ctx->shape.batch_chunk;
    t->topk_tile = ctx->shape.topk_tile;
    t->dense_chunk = ctx->shape.dense_chunk;
    t->visit_log = ctx->shape.visit_log;
    t->edge_records = ctx->shape.edge_records;
    return PTW_OK;
}

int ptw_tuning_set(ptw_ctx *ctx, const ptw_tuning *t)
{
    if (!ctx || !t) return fail(ctx, PTW_ERR_INVALID_ARG, "null argument");
    if (t->block < 64 || t->block > 1024 || t->block % 64)
        return fail(ctx, PTW_ERR_INVALID_ARG, "block must be a multiple of 64 in [64, 1024]");
    if (t->blocks_per_cu < 0 || t->blocks_per_cu > 32)
        return fail(ctx, PTW_ERR_INVALID_ARG, "blocks_per_cu must be in [0, 32]");
    if (t->chunk < 64) return fail(ctx, PTW_ERR_INVALID_ARG, "chunk must be >= 64");
    if (t->lds_hubs < 0 || t->lds_hubs > ptw::kMaxLdsHubs) return fail(ctx, PTW_ERR_INVALID_ARG, "lds_hubs must be in [0, 81408]");
    if (t->cache_slots != PTW_CACHE_AUTO &&
        (t->cache_slots < 0 || t->cache_slots > 16384 || (t->cache_slots & (t->cache_slots - 1))))
        return fail(ctx, PTW_ERR_INVALID_ARG, "cache_slots must be PTW_CACHE_AUTO, 0 or a power of two <= 16384");
    // u16 hub counters and the u32 cache share one workgroup's LDS (160 KiB per
    // CU, 1 KiB kept for the rest); the automatic setting may use 512 slots
    const int64_t slots = t->cache_slots == PTW_CACHE_AUTO ? ptw::kAutoCacheSlots : t->cache_slots;
    if ((((int64_t)t->lds_hubs + 1) / 2 + 2 * slots) * 4 > 159 * 1024)
        return fail(ctx, PTW_ERR_INVALID_ARG, "2 B per lds_hubs + 8 B per cache slot must fit in 159 KiB");
    if (t->batch_chunk < 64) return fail(ctx, PTW_ERR_INVALID_ARG, "batch_chunk must be >= 64");
    if (t->topk_tile < 256 || t->topk_tile > 4096 || (t->topk_tile & (t->topk_tile - 1)))
        return fail(ctx, PTW_ERR_INVALID_ARG, "topk_tile must be a power of two in [256, 4096]");
    if (t->dense_chunk < 0 || t->dense_chunk > (1 << 20) || t->dense_chunk % 1024)
        return fail(ctx, PTW_ERR_INVALID_ARG, "dense_chunk must be 0 or a multiple of 1024 up to 2^20");
    if (t->visit_log < 0 || (t->visit_log > 1 && t->visit_log % (int32_t)ptw::kVlogSeg))
        return fail(ctx, PTW_ERR_INVALID_ARG, "visit_log must be 0, 1 or a multiple of 1024 (entries)");
    if (t->edge_records < -1 || t->edge_records > 1)
        return fail(ctx, PTW_ERR_INVALID_ARG, "edge_records must be -1, 0 or 1");
    ptw::drop_batch_graph(ctx);
    ctx->shape.visit_log = t->visit_log;
    ctx->shape.edge_records = t->edge_records;
    ctx->shape.dense_chunk = t->dense_chunk;
    ctx->shape.batch_chunk = t->batch_chunk;
    ctx->shape.topk_tile = t->topk_tile;
    ctx->shape.block = t->block;
    ctx->shape.blocks_per_cu = t->blocks_per_cu;
    ctx->shape.chunk = t->chunk;
    ctx->shape.lds_hubs = t->lds_hubs;
    ctx->shape.cache_slots = t->cache_slots;
    return PTW_OK;
}

int ptw_buffer_alloc(ptw_ctx *ctx, uint64_t bytes, void **dptr)
{
    if (!ctx || !dptr) return fail(ctx, PTW_ERR_INVALID_ARG, "null argument");
    *dptr = nullptr;
    if (ctx->device < 0) return fail(ctx, PTW_ERR_UNSUPPORTED, "host-only context: no HIP device");
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    HIP_TRY(ctx, hipMalloc(dptr, bytes ? bytes : 4));
    return PTW_OK;
}

int ptw_buffer_free(ptw_ctx *ctx, void *dptr)
{
    if (!ctx) return PTW_ERR_INVALID_ARG;
    if (!dptr) return PTW_OK;
    if (ctx->device < 0) return fail(ctx, PTW_ERR_UNSUPPORTED, "host-only context: no HIP device");
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    HIP_TRY(ctx, hipFree(dptr));
    return PTW_OK;
}

int ptw_buffer_upload(ptw_ctx *ctx, void *dptr, const void *host, uint64_t bytes)
{
    if (!ctx || (bytes && (!dptr || !host))) return fail(ctx, PTW_ERR_INVALID_ARG, "null argument");
    if (ctx->device < 0) return fail(ctx, PTW_ERR_UNSUPPORTED, "host-only context: no HIP device");
    if (!bytes) return PTW_OK;
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    HIP_TRY(ctx, hipMemcpyAsync(dptr, host, bytes, hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return PTW_OK;
}

int ptw_buffer_download(ptw_ctx *ctx, const void *dptr, void *host, uint64_t bytes)
{
    if (!ctx || (bytes && (!dptr || !host))) return fail(ctx, PTW_ERR_INVALID_ARG, "null argument");
    if (ctx->device < 0) return fail(ctx, PTW_ERR_UNSUPPORTED, "host-only context: no HIP device");
    if (!bytes) return PTW_OK;
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    HIP_TRY(ctx, hipMemcpyAsync(host, dptr, bytes, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return PTW_OK;
}

int ptw_buffer_zero(ptw_ctx *ctx, void *dptr, uint64_t bytes)
{
    if (!ctx || (bytes && !dptr) || (bytes & 3)) return fail(ctx, PTW_ERR_INVALID_ARG, "bad buffer or size");
    if (ctx->device < 0) return fail(ctx, PTW_ERR_UNSUPPORTED, "host-only context: no HIP device");
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    HIP_TRY(ctx, ptw::launch_fill_u32((uint32_t *)dptr, bytes / 4, 0u, ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return PTW_OK;
}

int ptw_device_memory(ptw_ctx *ctx, uint64_t *free_bytes, uint64_t *total_bytes)
{
    if (!ctx) return PTW_ERR_INVALID_ARG;
    if (ctx->device < 0) return fail(ctx, PTW_ERR_UNSUPPORTED, "host-only context: no HIP device");
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    size_t f = 0, t = 0;
    HIP_TRY(ctx, hipMemGetInfo(&f, &t));
    if (free_bytes) *free_bytes = f;
    if (total_bytes) *total_bytes = t;
    return PTW_OK;
}

int ptw_bench_copy(ptw_ctx *ctx, uint64_t bytes, double *gb_per_s, double *kernel_ms)
{
    if (!ctx) return PTW_ERR_INVALID_ARG;
    if (ctx->device < 0) return fail(ctx, PTW_ERR_UNSUPPORTED, "host-only context: no HIP device");
    if (bytes == 0 || bytes % 16) return fail(ctx, PTW_ERR_INVALID_ARG, "copy size must be a positive multiple of 16");
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    struct Tmp {
        void *p = nullptr;
        ~Tmp() { if (p) (void)hipFree(p); }
    } a, b;
    HIP_TRY(ctx, hipMalloc(&a.p, bytes));
    HIP_TRY(ctx, hipMalloc(&b.p, bytes));
    HIP_TRY(ctx, ptw::launch_fill_u32((uint32_t *)a.p, bytes / 4, 0x5A5A5A5Au, ctx->stream));
    HIP_TRY(ctx, ptw::launch_copy_bench((const uint4 *)a.p, (uint4 *)b.p, bytes / 16, ctx->num_cus, ctx->stream));
    HIP_TRY(ctx, hipEventRecord(ctx->ev0, ctx->stream));
    HIP_TRY(ctx, ptw::launch_copy_bench((const uint4 *)a.p, (uint4 *)b.p, bytes / 16, ctx->num_cus, ctx->stream));
    HIP_TRY(ctx, hipEventRecord(ctx->ev1, ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    float ms = 0.f;
    HIP_TRY(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    if (kernel_ms) *kernel_ms = ms;
    if (gb_per_s) *gb_per_s = ms > 0 ? 2.0 * (double)bytes / (ms * 1e-3) / 1e9 : 0.0;
    return PTW_OK;
}

int ptw_bench_step(ptw_ctx *ctx, uint64_t steps, uint64_t seed, double *steps_per_s, double *kernel_ms)
{
    if (!ctx) return PTW_ERR_INVALID_ARG;
    if (!ctx->has_graph) return fail(ctx, PTW_ERR_NO_GRAPH, "no graph loaded");
    if (ctx->device < 0) return fail(ctx, PTW_ERR_UNSUPPORTED, "host-only context: no HIP device");
    if (steps == 0 || ctx->dev.E == 0) return fail(ctx, PTW_ERR_INVALID_ARG, "no steps or no edges");
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    struct Tmp {
        void *p = nullptr;
        ~Tmp() { if (p) (void)hipFree(p); }
    } starts, sink;
    HIP_TRY(ctx, hipMalloc(&starts.p, steps * 4));
    HIP_TRY(ctx, hipMalloc(&sink.p, 8));
    hipStream_t s = ctx->stream;
    const ptw::DeviceGraph &g = ctx->dev;
    HIP_TRY(ctx, ptw::launch_fill_u64((unsigned long long *)sink.p, 1, 0ull, s));
    HIP_TRY(ctx, ptw::launch_stationary_starts(g.row_ptr, g.V, g.E, steps, (uint32_t)seed, (uint32_t)(seed >> 32),
                                               (uint32_t *)starts.p, s));
    // warm-up launch over a slice, then the timed one
    HIP_TRY(ctx, ptw::launch_step_bench(g.row_ptr, g.col, (const uint32_t *)starts.p, steps / 8 + 1, (uint32_t)seed,
                                        (uint32_t)(seed >> 32) ^ 1u, (unsigned long long *)sink.p, ctx->num_cus, s));
    HIP_TRY(ctx, hipEventRecord(ctx->ev0, s));
    HIP_TRY(ctx, ptw::launch_step_bench(g.row_ptr, g.col, (const uint32_t *)starts.p, steps, (uint32_t)seed,
                                        (uint32_t)(seed >> 32), (unsigned long long *)sink.p, ctx->num_cus, s));
    HIP_TRY(ctx, hipEventRecord(ctx->ev1, s));
    HIP_TRY(ctx, hipStreamSynchronize(s));
    float ms = 0.f;
    HIP_TRY(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    if (kernel_ms) *kernel_ms = ms;
    if (steps_per_s) *steps_per_s = ms > 0 ? (double)steps / (ms * 1e-3) : 0.0;
    return PTW_OK;
}

int ptw_check_primitives(ptw_ctx *ctx, uint64_t n, int bits, uint64_t seed)
{
    if (!ctx) return PTW_ERR_INVALID_ARG;
    if (ctx->device < 0) return fail(ctx, PTW_ERR_UNSUPPORTED, "host-only context: no HIP device");
    if (n >= 0xFFFFFFFFull || bits < 1 || bits > 32)
        return fail(ctx, PTW_ERR_INVALID_ARG, "n must be below 2^32 - 1 and bits in 1..32");
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    try {
        std::string msg;
        const int r = ptw::prims::self_check(n, bits, seed, ctx->stream, msg);
        if (r == -2) return fail(ctx, PTW_ERR_HIP, msg);
        if (r != 0) return fail(ctx, PTW_ERR_FORMAT, msg);
        return PTW_OK;
    } catch (const std::bad_alloc &) {
        return fail(ctx, PTW_ERR_OOM, "host allocation failed in the primitive self-check");
    }
}

int ptw_bench_gather(ptw_ctx *ctx, uint64_t table_bytes, uint64_t gathers, uint64_t seed, double *gathers_per_s,
                     double *kernel_ms)
{
    if (!ctx) return PTW_ERR_INVALID_ARG;
    if (ctx->device < 0) return fail(ctx, PTW_ERR_UNSUPPORTED, "host-only context: no HIP device");
    if (table_bytes < 4 || gathers == 0) return fail(ctx, PTW_ERR_INVALID_ARG, "empty gather benchmark");
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    struct Tmp {
        void *p = nullptr;
        ~Tmp() { if (p) (void)hipFree(p); }
    } table, sink;
    const uint64_t words = table_bytes / 4;
    HIP_TRY(ctx, hipMalloc(&table.p, words * 4));
    HIP_TRY(ctx, hipMalloc(&sink.p, 8));
    HIP_TRY(ctx, ptw::launch_fill_u64((unsigned long long *)table.p, words / 2, 0x0123456789ABCDEFull, ctx->stream));
    HIP_TRY(ctx, ptw::launch_fill_u64((unsigned long long *)sink.p, 1, 0ull, ctx->stream));
    // warm-up launch, then the timed one
    HIP_TRY(ctx, ptw::launch_gather_bench((const uint32_t *)table.p, words, gathers / 4 + 1, (uint32_t)seed,
                                          (uint32_t)(seed >> 32) ^ 1u, (unsigned long long *)sink.p, ctx->num_cus,
                                          ctx->stream));
    HIP_TRY(ctx, hipEventRecord(ctx->ev0, ctx->stream));
    HIP_TRY(ctx, ptw::launch_gather_bench((const uint32_t *)table.p, words, gathers, (uint32_t)seed,
                                          (uint32_t)(seed >> 32), (unsigned long long *)sink.p, ctx->num_cus,
                                          ctx->stream));
    HIP_TRY(ctx, hipEventRecord(ctx->ev1, ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    float ms = 0.f;
    HIP_TRY(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    if (kernel_ms) *kernel_ms = ms;
    if (gathers_per_s) *gathers_per_s = ms > 0 ? (double)gathers / (ms * 1e-3) : 0.0;
    return PTW_OK;
}

}  // extern "C"
