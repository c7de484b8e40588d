// Budget predictor: a 6-64-64-1 ReLU MLP with BF16 weights and FP32
// accumulation, trained on the device (forward, backward, Adam).
//
// Parameters (FP32 master copy, flat):
//   W1[64][6] @0, b1[64] @384, W2[64][64] @448, b2[64] @4544, W3[64] @4608, b3 @4672
// The forward and backward passes read the weights rounded to BF16 (round to
// nearest even); biases stay FP32; every sum accumulates in FP32.
// Inputs are standardized on the device with the model's mean/std; the
// network predicts the standardized target.
//
// Gradients are deterministic: each 64-sample block writes its own partial
// gradient (samples summed in a fixed order), and a second kernel sums the
// partials over blocks in a fixed order.
#include "engine.hpp"

namespace ptw {

namespace {

constexpr int kIn = 6, kH = 64;
constexpr int oW1 = 0, oB1 = 384, oW2 = 448, oB2 = 4544, oW3 = 4608, oB3 = 4672;
constexpr int kBlock = 64;  // samples per backward block (one wavefront)

__device__ __forceinline__ float bf16_round(float x)
{
    uint32_t b = __float_as_uint(x);
    b = (b + 0x7FFFu + ((b >> 16) & 1u)) & 0xFFFF0000u;
    return __uint_as_float(b);
}

struct Weights {  // LDS copy, BF16-rounded weights as float, FP32 biases
    float w1[kH * kIn], b1[kH], w2[kH * kH], b2[kH], w3[kH], b3;
};

__device__ void load_weights(Weights &W, const float *params)
{
    for (int i = threadIdx.x; i < kH * kIn; i += blockDim.x) W.w1[i] = bf16_round(params[oW1 + i]);
    for (int i = threadIdx.x; i < kH * kH; i += blockDim.x) W.w2[i] = bf16_round(params[oW2 + i]);
    for (int i = threadIdx.x; i < kH; i += blockDim.x) {
        W.b1[i] = params[oB1 + i];
        W.b2[i] = params[oB2 + i];
        W.w3[i] = bf16_round(params[oW3 + i]);
    }
    if (threadIdx.x == 0) W.b3 = params[oB3];
    __syncthreads();
}

// norm: mean[6], std[6], y_mean, y_std
__device__ __forceinline__ void standardize(const float *x, const float *norm, float *xs)
{
#pragma unroll
    for (int k = 0; k < kIn; ++k) xs[k] = (x[k] - norm[k]) / norm[kIn + k];
}

__global__ __launch_bounds__(256) void mlp_infer_kernel(const float *params, const float *norm, const float *X,
                                                       uint32_t n, float *Y)
{
    __shared__ Weights W;
    load_weights(W, params);
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    float x[kIn];
    standardize(X + (size_t)s * kIn, norm, x);
    float h1[kH];
#pragma unroll
    for (int j = 0; j < kH; ++j) {
        float a = W.b1[j];
#pragma unroll
        for (int k = 0; k < kIn; ++k) a += W.w1[j * kIn + k] * x[k];
        h1[j] = a > 0.f ? a : 0.f;
    }
    float y = W.b3;
    for (int j = 0; j < kH; ++j) {
        float a = W.b2[j];
#pragma unroll
        for (int i = 0; i < kH; ++i) a += W.w2[j * kH + i] * h1[i];
        y += W.w3[j] * (a > 0.f ? a : 0.f);
    }
    Y[s] = y * norm[2 * kIn + 1] + norm[2 * kIn];  // de-standardize
}

// One block of kBlock samples: forward, backward, per-block partial gradient.
// idx selects the samples (a mini-batch); dy uses the mini-batch size `m`.
__global__ __launch_bounds__(kBlock) void mlp_grad_kernel(const float *params, const float *norm, const float *X,
                                                          const float *T, const uint32_t *idx, uint32_t m,
                                                          float *partial, float *loss_partial)
{
    __shared__ Weights W;
    __shared__ float s_h1[kBlock][kH + 1], s_h2[kBlock][kH + 1], s_d1[kBlock][kH + 1], s_d2[kBlock][kH + 1];
    __shared__ float s_x[kBlock][kIn], s_dy[kBlock];
    load_weights(W, params);
    const uint32_t t = threadIdx.x;
    const uint32_t r = blockIdx.x * kBlock + t;  // row of the mini-batch
    const bool live = r < m;
    float x[kIn] = {0, 0, 0, 0, 0, 0};
    float dy = 0.f;
    float z1[kH], z2[kH];
    if (live) {
        const uint32_t sidx = idx ? idx[r] : r;
        standardize(X + (size_t)sidx * kIn, norm, x);
        const float target = (T[sidx] - norm[2 * kIn]) / norm[2 * kIn + 1];
#pragma unroll
        for (int j = 0; j < kH; ++j) {
            float a = W.b1[j];
#pragma unroll
            for (int k = 0; k < kIn; ++k) a += W.w1[j * kIn + k] * x[k];
            z1[j] = a;
        }
        float y = W.b3;
        for (int j = 0; j < kH; ++j) {
            float a = W.b2[j];
#pragma unroll
            for (int i = 0; i < kH; ++i) a += W.w2[j * kH + i] * (z1[i] > 0.f ? z1[i] : 0.f);
            z2[j] = a;
            y += W.w3[j] * (a > 0.f ? a : 0.f);
        }
        const float e = y - target;
        dy = 2.f * e / (float)m;
        s_dy[t] = dy;
        loss_partial[blockIdx.x * kBlock + t] = e * e / (float)m;
        for (int j = 0; j < kH; ++j) {
            s_h1[t][j] = z1[j] > 0.f ? z1[j] : 0.f;
            s_h2[t][j] = z2[j] > 0.f ? z2[j] : 0.f;
            s_d2[t][j] = z2[j] > 0.f ? dy * W.w3[j] : 0.f;
        }
        for (int i = 0; i < kH; ++i) {
            float a = 0.f;
            for (int j = 0; j < kH; ++j) a += s_d2[t][j] * W.w2[j * kH + i];
            s_d1[t][i] = z1[i] > 0.f ? a : 0.f;
        }
#pragma unroll
        for (int k = 0; k < kIn; ++k) s_x[t][k] = x[k];
    } else {
        s_dy[t] = 0.f;
        loss_partial[blockIdx.x * kBlock + t] = 0.f;
        for (int j = 0; j < kH; ++j) {
            s_h1[t][j] = 0.f;
            s_h2[t][j] = 0.f;
            s_d2[t][j] = 0.f;
            s_d1[t][j] = 0.f;
        }
#pragma unroll
        for (int k = 0; k < kIn; ++k) s_x[t][k] = 0.f;
    }
    __syncthreads();
    // partial gradient of this block: samples summed in index order
    float *g = partial + (size_t)blockIdx.x * (oB3 + 1);
    for (int p = t; p <= oB3; p += kBlock) {
        float a = 0.f;
        if (p < oB1) {
            const int j = p / kIn, k = p % kIn;
            for (int s = 0; s < kBlock; ++s) a += s_d1[s][j] * s_x[s][k];
        } else if (p < oW2) {
            const int j = p - oB1;
            for (int s = 0; s < kBlock; ++s) a += s_d1[s][j];
        } else if (p < oB2) {
            const int j = (p - oW2) / kH, i = (p - oW2) % kH;
            for (int s = 0; s < kBlock; ++s) a += s_d2[s][j] * s_h1[s][i];
        } else if (p < oW3) {
            const int j = p - oB2;
            for (int s = 0; s < kBlock; ++s) a += s_d2[s][j];
        } else if (p < oB3) {
            const int j = p - oW3;
            for (int s = 0; s < kBlock; ++s) a += s_dy[s] * s_h2[s][j];
        } else {
            for (int s = 0; s < kBlock; ++s) a += s_dy[s];
        }
        g[p] = a;
    }
}

__global__ void mlp_reduce_kernel(const float *partial, uint32_t blocks, float *grad, const float *loss_partial,
                                  uint32_t rows, float *loss)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p <= oB3) {
        float a = 0.f;
        for (uint32_t b = 0; b < blocks; ++b) a += partial[(size_t)b * (oB3 + 1) + p];
        grad[p] = a;
    }
    if (p == 0 && loss) {
        float a = 0.f;
        for (uint32_t i = 0; i < rows; ++i) a += loss_partial[i];
        *loss = a;
    }
}

__global__ void mlp_adam_kernel(float *params, float *m, float *v, const float *grad, float lr, float b1, float b2,
                                float eps, float c1, float c2)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p > oB3) return;
    const float g = grad[p];
    const float mm = b1 * m[p] + (1.f - b1) * g;
    const float vv = b2 * v[p] + (1.f - b2) * g * g;
    m[p] = mm;
    v[p] = vv;
    params[p] -= lr * (mm / c1) / (sqrtf(vv / c2) + eps);
}

}  // namespace

int mlp_num_params() { return oB3 + 1; }

hipError_t launch_mlp_infer(const float *params, const float *norm, const float *X, uint32_t n, float *Y,
                            hipStream_t s)
{
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(mlp_infer_kernel, dim3((n + 255) / 256), dim3(256), 0, s, params, norm, X, n, Y);
    return hipGetLastError();
}

hipError_t launch_mlp_grad(const float *params, const float *norm, const float *X, const float *T,
                           const uint32_t *idx, uint32_t m, float *partial, float *loss_partial, float *grad,
                           float *loss, hipStream_t s)
{
    const uint32_t blocks = (m + kBlock - 1) / kBlock;
    hipLaunchKernelGGL(mlp_grad_kernel, dim3(blocks), dim3(kBlock), 0, s, params, norm, X, T, idx, m, partial,
                       loss_partial);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(mlp_reduce_kernel, dim3((oB3 + 1 + 255) / 256), dim3(256), 0, s, partial, blocks, grad,
                       loss_partial, blocks * kBlock, loss);
    return hipGetLastError();
}

hipError_t launch_mlp_adam(float *params, float *m, float *v, const float *grad, float lr, int step, hipStream_t s)
{
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.f - powf(b1, (float)step), c2 = 1.f - powf(b2, (float)step);
    hipLaunchKernelGGL(mlp_adam_kernel, dim3((oB3 + 1 + 255) / 256), dim3(256), 0, s, params, m, v, grad, lr, b1, b2,
                       1e-8f, c1, c2);
    return hipGetLastError();
}

}  // namespace ptw
