// FP64 power iteration on the device for a block of B sources at once: the
// expected visit vector per walker of single-source walks,
//   x = e_s + sum_{i>=1} P(L >= i) e_s P^i,
// P(L >= i) = (1-tp)^(i-1) under reference semantics, (1-tp)^i textbook.
// Pull form over the transpose CSR: y'[v] = sum_{u -> v} y[u] / deg(u).
// Rows hold the B sources contiguously (B doubles per vertex), so every
// neighbour read of a wavefront is one contiguous B x 8-byte row.
//
// In-edge lists are cut into pieces of at most kPiece edges so that a hub
// with 10^5 in-edges is spread over many wavefronts: pass 1 sums each piece
// into a partial row, pass 2 adds a vertex's partial rows in piece order.
// Both passes are deterministic.
#include "../../include/ptw.h"
#include "engine.hpp"

namespace ptw {

namespace {

__global__ __launch_bounds__(256) void exact_piece_kernel(const uint32_t *t_col, const uint32_t *out_deg,
                                                         const uint32_t *piece_begin, uint64_t pieces, uint32_t B,
                                                         const double *y, double *partial)
{
    const uint32_t lane = threadIdx.x & 63;
    const uint64_t wave = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const uint64_t waves = ((uint64_t)gridDim.x * blockDim.x) >> 6;
    for (uint64_t pc = wave; pc < pieces; pc += waves) {
        if (lane >= B) continue;
        const uint32_t b = piece_begin[pc], e = piece_begin[pc + 1];
        double acc = 0.0;
        for (uint32_t k = b; k < e; ++k) {
            const uint32_t u = t_col[k];
            acc += y[(uint64_t)u * B + lane] / (double)out_deg[u];
        }
        partial[pc * B + lane] = acc;
    }
}

__global__ __launch_bounds__(256) void exact_vertex_kernel(const uint32_t *vertex_piece, uint32_t V, uint32_t B,
                                                          const double *partial, double *y_next, double *x,
                                                          const double *weight)
{
    const uint32_t lane = threadIdx.x & 63;
    const uint64_t wave = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const uint64_t waves = ((uint64_t)gridDim.x * blockDim.x) >> 6;
    if (lane >= B) return;
    const double wgt = weight[lane];
    for (uint64_t v = wave; v < V; v += waves) {
        double acc = 0.0;
        for (uint32_t pc = vertex_piece[v]; pc < vertex_piece[v + 1]; ++pc) acc += partial[(uint64_t)pc * B + lane];
        y_next[v * B + lane] = acc;
        x[v * B + lane] += wgt * acc;
    }
}

__global__ void exact_init_kernel(uint32_t V, uint32_t B, const uint32_t *src_dev, double *y, double *x)
{
    const uint64_t n = (uint64_t)V * B;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const uint32_t v = (uint32_t)(i / B), b = (uint32_t)(i % B);
        const double e = src_dev[b] == v ? 1.0 : 0.0;
        y[i] = e;
        x[i] = e;
    }
}

// out[b * V + g] = x[perm[g] * B + b]  (graph ids, one row per source)
__global__ void exact_export_kernel(uint32_t V, uint32_t B, const uint32_t *perm, const double *x, double *out)
{
    const uint64_t n = (uint64_t)V * B;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const uint32_t b = (uint32_t)(i / V), g = (uint32_t)(i % V);
        out[i] = x[(uint64_t)(perm ? perm[g] : g) * B + b];
    }
}

}  // namespace

hipError_t launch_exact_init(uint32_t V, uint32_t B, const uint32_t *src_dev, double *y, double *x, hipStream_t s)
{
    hipLaunchKernelGGL(exact_init_kernel, dim3(4096), dim3(256), 0, s, V, B, src_dev, y, x);
    return hipGetLastError();
}

hipError_t launch_exact_step(const ExactGraph &g, uint32_t B, const double *y, double *y_next, double *x,
                             double *partial, const double *weight, int num_cus, hipStream_t s)
{
    hipLaunchKernelGGL(exact_piece_kernel, dim3(num_cus * 32), dim3(256), 0, s, g.t_col, g.out_deg, g.piece_begin,
                       g.pieces, B, y, partial);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(exact_vertex_kernel, dim3(num_cus * 32), dim3(256), 0, s, g.vertex_piece, g.V, B, partial,
                       y_next, x, weight);
    return hipGetLastError();
}

hipError_t launch_exact_export(uint32_t V, uint32_t B, const uint32_t *perm, const double *x, double *out,
                               hipStream_t s)
{
    hipLaunchKernelGGL(exact_export_kernel, dim3(4096), dim3(256), 0, s, V, B, perm, x, out);
    return hipGetLastError();
}

}  // namespace ptw
