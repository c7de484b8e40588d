// C ABI of the device FP64 solver (ptw_ppr_exact): the reference process's
// expected visit vectors for blocks of up to 64 sources, used to audit walk
// estimates and to label budgets for the learned predictor.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/ptw.h"
#include "ctx.hpp"
#include "engine.hpp"

namespace ptw {

namespace {

// Transpose of the device-id CSR, cut into pieces of at most kExactPiece
// in-edges, plus out-degrees; built once per graph.
int ensure_transpose(ptw_ctx *ctx)
{
    DeviceGraph &d = ctx->dev;
    if (d.t_col) return PTW_OK;
    const uint32_t V = d.V;
    std::vector<uint32_t> row(V + 1), col;
    // device-id CSR on the host
    const HostCsr &h = ctx->host;
    auto dev_of = [&](uint32_t g) { return d.relabelled ? ctx->h_perm[g] : g; };
    std::vector<uint32_t> indeg(V, 0), outdeg(V, 0);
    for (uint32_t g = 0; g < V; ++g) {
        const uint32_t u = dev_of(g);
        outdeg[u] = h.row_ptr[g + 1] - h.row_ptr[g];
        for (uint32_t k = h.row_ptr[g]; k < h.row_ptr[g + 1]; ++k) indeg[dev_of(h.col[k])]++;
    }
    row[0] = 0;
    for (uint32_t v = 0; v < V; ++v) row[v + 1] = row[v] + indeg[v];
    col.resize(h.E());
    std::vector<uint32_t> fill(row.begin(), row.end() - 1);
    for (uint32_t g = 0; g < V; ++g) {
        const uint32_t u = dev_of(g);
        for (uint32_t k = h.row_ptr[g]; k < h.row_ptr[g + 1]; ++k) col[fill[dev_of(h.col[k])]++] = u;
    }
    std::vector<uint32_t> piece_begin, vertex_piece(V + 1);
    for (uint32_t v = 0; v < V; ++v) {
        vertex_piece[v] = (uint32_t)piece_begin.size();
        for (uint32_t b = row[v]; b < row[v + 1]; b += kExactPiece) piece_begin.push_back(b);
    }
    vertex_piece[V] = (uint32_t)piece_begin.size();
    d.pieces = piece_begin.size();
    piece_begin.push_back(row[V]);
    int rc;
    if ((rc = upload(ctx, &d.t_col, col.data(), col.size()))) return rc;
    if ((rc = upload(ctx, &d.out_deg, outdeg.data(), outdeg.size()))) return rc;
    if ((rc = upload(ctx, &d.piece_begin, piece_begin.data(), piece_begin.size()))) return rc;
    if ((rc = upload(ctx, &d.vertex_piece, vertex_piece.data(), vertex_piece.size()))) return rc;
    return PTW_OK;
}

}  // namespace

}  // namespace ptw

using ptw::fail;

extern "C" {

int ptw_ppr_exact(ptw_ctx *ctx, const uint32_t *sources, const double *tps, uint32_t n, int32_t semantics,
                  double tol, double *out, uint32_t k, uint32_t *topk_vertex, double *topk_value,
                  uint32_t *iterations)
{
    if (!ctx) return PTW_ERR_INVALID_ARG;
    if (!ctx->has_graph) return fail(ctx, PTW_ERR_NO_GRAPH, "no graph loaded");
    if (ctx->device < 0) return fail(ctx, PTW_ERR_UNSUPPORTED, "host-only context: the solver needs a HIP device");
    if (n && (!sources || !tps)) return fail(ctx, PTW_ERR_INVALID_ARG, "null sources or tps");
    if (semantics != PTW_SEM_REFERENCE && semantics != PTW_SEM_TEXTBOOK)
        return fail(ctx, PTW_ERR_INVALID_ARG, "unknown semantics");
    if (!(tol > 0.0)) return fail(ctx, PTW_ERR_INVALID_ARG, "tol must be > 0");
    if (k && (!topk_vertex || !topk_value)) return fail(ctx, PTW_ERR_INVALID_ARG, "top-k outputs required");
    const uint32_t V = ctx->dev.V;
    if (k > V) return fail(ctx, PTW_ERR_INVALID_ARG, "k exceeds the vertex count");
    for (uint32_t i = 0; i < n; ++i) {
        if (sources[i] >= V) return fail(ctx, PTW_ERR_INVALID_ARG, "source out of range");
        if (!(tps[i] > 0.0 && tps[i] <= 1.0)) return fail(ctx, PTW_ERR_INVALID_ARG, "tp must be in (0, 1]");
    }
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    int rc = ptw::ensure_transpose(ctx);
    if (rc) return rc;
    hipStream_t s = ctx->stream;
    const uint32_t BMAX = 64;
    struct Tmp {
        void *p = nullptr;
        ~Tmp() { if (p) (void)hipFree(p); }
    } ty, tz, tx, tw, ts, to, tp_;
    const uint32_t Bcap = std::min(n, BMAX);
    if (n == 0) return PTW_OK;
    const ptw::ExactGraph eg{V, ctx->dev.t_col, ctx->dev.out_deg, ctx->dev.piece_begin, ctx->dev.vertex_piece,
                             ctx->dev.pieces};
    HIP_TRY(ctx, hipMalloc(&tp_.p, (size_t)std::max<uint64_t>(ctx->dev.pieces, 1) * Bcap * 8));
    HIP_TRY(ctx, hipMalloc(&ty.p, (size_t)V * Bcap * 8));
    HIP_TRY(ctx, hipMalloc(&tz.p, (size_t)V * Bcap * 8));
    HIP_TRY(ctx, hipMalloc(&tx.p, (size_t)V * Bcap * 8));
    HIP_TRY(ctx, hipMalloc(&ts.p, Bcap * 4));
    HIP_TRY(ctx, hipMalloc(&to.p, (size_t)V * Bcap * 8));
    std::vector<double> host((size_t)V * Bcap);
    uint32_t max_iter_used = 0;
    for (uint32_t b0 = 0; b0 < n; b0 += BMAX) {
        const uint32_t B = std::min(BMAX, n - b0);
        // iterations: remaining mass after i steps <= (1 - tp)^i / tp
        double tpmin = 1.0;
        for (uint32_t b = 0; b < B; ++b) tpmin = std::min(tpmin, tps[b0 + b]);
        uint32_t iters = 1;
        while (iters < 100000 && std::pow(1.0 - tpmin, (double)iters) / tpmin >= tol) ++iters;
        if (tpmin >= 1.0) iters = 1;
        max_iter_used = std::max(max_iter_used, iters);
        // per-iteration weights P(L >= i), i = 1..iters
        std::vector<double> w((size_t)iters * B);
        for (uint32_t b = 0; b < B; ++b) {
            const double q = 1.0 - tps[b0 + b];
            double pw = semantics == PTW_SEM_TEXTBOOK ? q : 1.0;
            for (uint32_t i = 0; i < iters; ++i) {
                w[(size_t)i * B + b] = pw;
                pw *= q;
            }
        }
        if (tw.p) (void)hipFree(tw.p);
        tw.p = nullptr;
        HIP_TRY(ctx, hipMalloc(&tw.p, w.size() * 8));
        HIP_TRY(ctx, hipMemcpyAsync(tw.p, w.data(), w.size() * 8, hipMemcpyHostToDevice, s));
        std::vector<uint32_t> src(B);
        for (uint32_t b = 0; b < B; ++b) src[b] = ctx->dev.relabelled ? ctx->h_perm[sources[b0 + b]] : sources[b0 + b];
        HIP_TRY(ctx, hipMemcpyAsync(ts.p, src.data(), B * 4, hipMemcpyHostToDevice, s));
        double *y = (double *)ty.p, *z = (double *)tz.p, *x = (double *)tx.p;
        HIP_TRY(ctx, ptw::launch_exact_init(V, B, (const uint32_t *)ts.p, y, x, s));
        for (uint32_t i = 0; i < iters; ++i) {
            HIP_TRY(ctx, ptw::launch_exact_step(eg, B, y, z, x, (double *)tp_.p, (const double *)tw.p + (size_t)i * B,
                                                ctx->num_cus, s));
            std::swap(y, z);
        }
        HIP_TRY(ctx, ptw::launch_exact_export(V, B, ctx->dev.perm, x, (double *)to.p, s));
        HIP_TRY(ctx, hipMemcpyAsync(host.data(), to.p, (size_t)V * B * 8, hipMemcpyDeviceToHost, s));
        HIP_TRY(ctx, hipStreamSynchronize(s));
        if (out) std::copy(host.begin(), host.begin() + (size_t)V * B, out + (size_t)b0 * V);
        if (k) {
            ptw::parallel_for(B, [&](uint64_t lo, uint64_t hi) {
                std::vector<uint32_t> idx(V);
                for (uint64_t b = lo; b < hi; ++b) {
                    const double *xb = host.data() + b * V;
                    std::iota(idx.begin(), idx.end(), 0u);
                    auto better = [xb](uint32_t a, uint32_t c) { return xb[a] > xb[c] || (xb[a] == xb[c] && a < c); };
                    std::partial_sort(idx.begin(), idx.begin() + k, idx.end(), better);
                    for (uint32_t j = 0; j < k; ++j) {
                        topk_vertex[(b0 + b) * k + j] = idx[j];
                        topk_value[(b0 + b) * k + j] = xb[idx[j]];
                    }
                }
            }, 16, 1);
        }
    }
    if (iterations) *iterations = max_iter_used;
    return PTW_OK;
}

}  // extern "C"
