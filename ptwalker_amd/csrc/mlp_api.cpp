// C ABI of the budget predictor (ptw_mlp_*): a 6-64-64-1 MLP whose forward,
// backward and Adam steps run as HIP kernels (mlp.hip).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <new>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/ptw.h"
#include "ctx.hpp"
#include "engine.hpp"
#include "philox.hpp"

namespace ptw {
int mlp_num_params();
hipError_t launch_mlp_infer(const float *params, const float *norm, const float *X, uint32_t n, float *Y,
                            hipStream_t s);
hipError_t launch_mlp_grad(const float *params, const float *norm, const float *X, const float *T,
                           const uint32_t *idx, uint32_t m, float *partial, float *loss_partial, float *grad,
                           float *loss, hipStream_t s);
hipError_t launch_mlp_adam(float *params, float *m, float *v, const float *grad, float lr, int step, hipStream_t s);
}  // namespace ptw

struct ptw_mlp {
    ptw_ctx *ctx = nullptr;
    float *params = nullptr, *m = nullptr, *v = nullptr, *grad = nullptr, *norm = nullptr;
    float *partial = nullptr, *loss_partial = nullptr, *loss = nullptr;
    uint32_t max_rows = 0;  // capacity of partial / loss_partial (rows of a gradient call)
    // inference inputs and outputs, grown and kept: a forward call frees
    // nothing, so it never waits for other streams' work (a free does)
    float *fx = nullptr, *fy = nullptr;
    uint32_t f_cap = 0;
    float h_norm[14] = {0, 0, 0, 0, 0, 0, 1, 1, 1, 1, 1, 1, 0, 1};
    int step = 0;
};

namespace {

using ptw::fail;

constexpr uint32_t kRowsPerBlock = 64;

int ensure_rows(ptw_mlp *M, uint32_t rows)
{
    ptw_ctx *ctx = M->ctx;
    if (rows <= M->max_rows) return PTW_OK;
    const uint32_t cap = std::max<uint32_t>(rows, 4096);
    const uint32_t blocks = (cap + kRowsPerBlock - 1) / kRowsPerBlock;
    if (M->partial) (void)hipFree(M->partial);
    if (M->loss_partial) (void)hipFree(M->loss_partial);
    M->partial = M->loss_partial = nullptr;
    HIP_TRY(ctx, hipMalloc((void **)&M->partial, (size_t)blocks * ptw::mlp_num_params() * 4));
    HIP_TRY(ctx, hipMalloc((void **)&M->loss_partial, (size_t)blocks * kRowsPerBlock * 4));
    M->max_rows = blocks * kRowsPerBlock;
    return PTW_OK;
}

struct DevBuf {
    void *p = nullptr;
    ~DevBuf() { if (p) (void)hipFree(p); }
};

}  // namespace

extern "C" {

int ptw_mlp_create(ptw_ctx *ctx, uint64_t seed, ptw_mlp **out)
{
    if (!ctx || !out) return fail(ctx, PTW_ERR_INVALID_ARG, "null argument");
    if (ctx->device < 0) return fail(ctx, PTW_ERR_UNSUPPORTED, "host-only context: the MLP runs on the device");
    *out = nullptr;
    ptw_mlp *M = new (std::nothrow) ptw_mlp;
    if (!M) return PTW_ERR_OOM;
    M->ctx = ctx;
    const int P = ptw::mlp_num_params();
    // He-uniform init from Philox(seed, parameter index); biases zero
    std::vector<float> p(P, 0.f);
    auto uni = [&](int i) {
        ptw::u32x4 r = ptw::philox4x32_10((uint32_t)i, 0, 0, 0x4D4C5031u, (uint32_t)seed, (uint32_t)(seed >> 32));
        return ((float)(r.x >> 8) + 0.5f) / 16777216.0f * 2.f - 1.f;
    };
    const float a1 = std::sqrt(6.f / 6.f), a2 = std::sqrt(6.f / 64.f), a3 = std::sqrt(6.f / 64.f);
    for (int i = 0; i < 384; ++i) p[i] = a1 * uni(i);
    for (int i = 448; i < 4544; ++i) p[i] = a2 * uni(i);
    for (int i = 4608; i < 4672; ++i) p[i] = a3 * uni(i);
    int rc = PTW_OK;
    auto cleanup = [&]() {
        ptw_mlp_destroy(M);
        return rc;
    };
    if (hipSetDevice(ctx->device) != hipSuccess) return rc = fail(ctx, PTW_ERR_HIP, "hipSetDevice failed"), cleanup();
    for (float **b : {&M->params, &M->m, &M->v, &M->grad}) {
        if (hipMalloc((void **)b, (size_t)P * 4) != hipSuccess) {
            rc = fail(ctx, PTW_ERR_OOM, "MLP allocation failed");
            return cleanup();
        }
    }
    if (hipMalloc((void **)&M->norm, 14 * 4) != hipSuccess || hipMalloc((void **)&M->loss, 4) != hipSuccess) {
        rc = fail(ctx, PTW_ERR_OOM, "MLP allocation failed");
        return cleanup();
    }
    std::vector<float> zeros(P, 0.f);
    if (hipMemcpy(M->params, p.data(), (size_t)P * 4, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(M->m, zeros.data(), (size_t)P * 4, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(M->v, zeros.data(), (size_t)P * 4, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(M->norm, M->h_norm, 14 * 4, hipMemcpyHostToDevice) != hipSuccess) {
        rc = fail(ctx, PTW_ERR_HIP, "MLP upload failed");
        return cleanup();
    }
    *out = M;
    return PTW_OK;
}

void ptw_mlp_destroy(ptw_mlp *M)
{
    if (!M) return;
    if (M->ctx && M->ctx->stream) (void)hipStreamSynchronize(M->ctx->stream);
    for (float *b : {M->params, M->m, M->v, M->grad, M->norm, M->partial, M->loss_partial, M->loss, M->fx, M->fy})
        if (b) (void)hipFree(b);
    delete M;
}

int ptw_mlp_num_params(void) { return ptw::mlp_num_params(); }

int ptw_mlp_set_normalization(ptw_mlp *M, const float *mean, const float *stdev, float y_mean, float y_std)
{
    if (!M || !mean || !stdev) return fail(M ? M->ctx : nullptr, PTW_ERR_INVALID_ARG, "null argument");
    for (int k = 0; k < 6; ++k)
        if (!(stdev[k] > 0.f) || !std::isfinite(mean[k]))
            return fail(M->ctx, PTW_ERR_INVALID_ARG, "feature std must be > 0 and means finite");
    if (!(y_std > 0.f)) return fail(M->ctx, PTW_ERR_INVALID_ARG, "target std must be > 0");
    for (int k = 0; k < 6; ++k) {
        M->h_norm[k] = mean[k];
        M->h_norm[6 + k] = stdev[k];
    }
    M->h_norm[12] = y_mean;
    M->h_norm[13] = y_std;
    HIP_TRY(M->ctx, hipMemcpy(M->norm, M->h_norm, 14 * 4, hipMemcpyHostToDevice));
    return PTW_OK;
}

int ptw_mlp_get_params(ptw_mlp *M, float *params, float *norm)
{
    if (!M) return PTW_ERR_INVALID_ARG;
    if (params) HIP_TRY(M->ctx, hipMemcpy(params, M->params, (size_t)ptw::mlp_num_params() * 4, hipMemcpyDeviceToHost));
    if (norm) std::copy(M->h_norm, M->h_norm + 14, norm);
    return PTW_OK;
}

int ptw_mlp_set_params(ptw_mlp *M, const float *params)
{
    if (!M || !params) return fail(M ? M->ctx : nullptr, PTW_ERR_INVALID_ARG, "null argument");
    HIP_TRY(M->ctx, hipMemcpy(M->params, params, (size_t)ptw::mlp_num_params() * 4, hipMemcpyHostToDevice));
    return PTW_OK;
}

int ptw_mlp_forward(ptw_mlp *M, const float *x, uint32_t n, float *y)
{
    if (!M || (n && (!x || !y))) return fail(M ? M->ctx : nullptr, PTW_ERR_INVALID_ARG, "null argument");
    if (n == 0) return PTW_OK;
    ptw_ctx *ctx = M->ctx;
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    if (n > M->f_cap) {
        HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
        if (M->fx) (void)hipFree(M->fx);
        if (M->fy) (void)hipFree(M->fy);
        M->fx = M->fy = nullptr;
        M->f_cap = 0;
        const uint32_t cap = std::max<uint32_t>(n, 1024);
        HIP_TRY(ctx, hipMalloc((void **)&M->fx, (size_t)cap * 6 * 4));
        HIP_TRY(ctx, hipMalloc((void **)&M->fy, (size_t)cap * 4));
        M->f_cap = cap;
    }
    HIP_TRY(ctx, hipMemcpyAsync(M->fx, x, (size_t)n * 24, hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(ctx, ptw::launch_mlp_infer(M->params, M->norm, M->fx, n, M->fy, ctx->stream));
    HIP_TRY(ctx, hipMemcpyAsync(y, M->fy, (size_t)n * 4, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return PTW_OK;
}

int ptw_mlp_gradients(ptw_mlp *M, const float *x, const float *t, uint32_t n, float *grad, float *loss)
{
    if (!M || !x || !t || n == 0) return fail(M ? M->ctx : nullptr, PTW_ERR_INVALID_ARG, "bad arguments");
    ptw_ctx *ctx = M->ctx;
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    int rc = ensure_rows(M, n);
    if (rc) return rc;
    DevBuf dx, dt;
    HIP_TRY(ctx, hipMalloc(&dx.p, (size_t)n * 24));
    HIP_TRY(ctx, hipMalloc(&dt.p, (size_t)n * 4));
    HIP_TRY(ctx, hipMemcpyAsync(dx.p, x, (size_t)n * 24, hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(ctx, hipMemcpyAsync(dt.p, t, (size_t)n * 4, hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(ctx, ptw::launch_mlp_grad(M->params, M->norm, (const float *)dx.p, (const float *)dt.p, nullptr, n,
                                      M->partial, M->loss_partial, M->grad, M->loss, ctx->stream));
    if (grad)
        HIP_TRY(ctx, hipMemcpyAsync(grad, M->grad, (size_t)ptw::mlp_num_params() * 4, hipMemcpyDeviceToHost,
                                    ctx->stream));
    float l = 0.f;
    HIP_TRY(ctx, hipMemcpyAsync(&l, M->loss, 4, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    if (loss) *loss = l;
    return PTW_OK;
}

int ptw_mlp_train(ptw_mlp *M, const float *x, const float *t, uint32_t n, uint32_t epochs, uint32_t batch, float lr,
                  uint64_t seed, float *final_loss)
{
    if (!M || !x || !t || n == 0 || batch == 0 || !(lr > 0.f))
        return fail(M ? M->ctx : nullptr, PTW_ERR_INVALID_ARG, "bad arguments");
    ptw_ctx *ctx = M->ctx;
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    batch = std::min(batch, n);
    int rc = ensure_rows(M, batch);
    if (rc) return rc;
    DevBuf dx, dt, di;
    HIP_TRY(ctx, hipMalloc(&dx.p, (size_t)n * 24));
    HIP_TRY(ctx, hipMalloc(&dt.p, (size_t)n * 4));
    HIP_TRY(ctx, hipMalloc(&di.p, (size_t)n * 4));
    hipStream_t s = ctx->stream;
    HIP_TRY(ctx, hipMemcpyAsync(dx.p, x, (size_t)n * 24, hipMemcpyHostToDevice, s));
    HIP_TRY(ctx, hipMemcpyAsync(dt.p, t, (size_t)n * 4, hipMemcpyHostToDevice, s));
    std::vector<uint32_t> order(n);
    std::iota(order.begin(), order.end(), 0u);
    float last = 0.f;
    for (uint32_t ep = 0; ep < epochs; ++ep) {
        // Fisher-Yates shuffle driven by Philox(seed, epoch, position)
        for (uint32_t i = n - 1; i > 0; --i) {
            ptw::u32x4 r = ptw::philox4x32_10(i, ep, 0, 0x53485546u, (uint32_t)seed, (uint32_t)(seed >> 32));
            std::swap(order[i], order[ptw::mulhi64(r.x, r.y, i + 1)]);
        }
        HIP_TRY(ctx, hipMemcpyAsync(di.p, order.data(), (size_t)n * 4, hipMemcpyHostToDevice, s));
        for (uint32_t b0 = 0; b0 + batch <= n; b0 += batch) {
            HIP_TRY(ctx, ptw::launch_mlp_grad(M->params, M->norm, (const float *)dx.p, (const float *)dt.p,
                                              (const uint32_t *)di.p + b0, batch, M->partial, M->loss_partial,
                                              M->grad, M->loss, s));
            HIP_TRY(ctx, ptw::launch_mlp_adam(M->params, M->m, M->v, M->grad, lr, ++M->step, s));
        }
        HIP_TRY(ctx, hipMemcpyAsync(&last, M->loss, 4, hipMemcpyDeviceToHost, s));
        HIP_TRY(ctx, hipStreamSynchronize(s));  // order[] is rewritten next epoch
    }
    if (final_loss) *final_loss = last;
    return PTW_OK;
}

}  // extern "C"
