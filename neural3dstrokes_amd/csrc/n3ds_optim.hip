// Fused stroke-parameter update: gradient sanitation, Adam and the per-column
// parameter clamp in one pass over each parameter tensor.
//
// Semantics of the reference's training step for the stroke field:
//  * clip_gradients (source/utils/training.py:67-77): optional global-norm
//    clipping (accelerator.clip_grad_norm_), optional value clipping
//    (clip_grad_value_), then grad.nan_to_num_();
//  * create_optimizer (training.py:119-136): torch.optim.Adam with betas
//    (0.9, 0.99), eps 1e-15 (source/configs.py:82-84), whose update is
//    m = lerp(m, g, 1 - b1); v = v * b2 + (1 - b2) g^2;
//    p -= lr / (1 - b1^t) * m / (sqrt(v) / sqrt(1 - b2^t) + eps);
//  * StrokeField.clip_params (source/models.py:576-584): per-column clamp of the
//    shape / colour parameters to their ranges, density >= 0.
// Memory-bound elementwise work (32 bytes per parameter: p, g, m, v read and
// written); one thread per element, the column bounds in LDS.
#include "n3ds_capi.h"

namespace n3ds {

constexpr int OPT_MAX_COLS = 1024;

__device__ __forceinline__ float nan_to_num(float g) {
    if (isnan(g)) return 0.f;
    if (isinf(g)) return g > 0.f ? 3.40282347e38f : -3.40282347e38f;
    return g;
}

// torch.clamp: min(max(x, lo), hi) with NaN passed through (fminf/fmaxf would drop it)
__device__ __forceinline__ float torch_clamp(float x, float lo, float hi) {
    const float y = x < lo ? lo : x;
    return y > hi ? hi : y;
}

// torch.lerp for float: weight < 0.5 ? a + w (b - a) : b - (b - a) (1 - w)
__device__ __forceinline__ float torch_lerp(float a, float b, float w) {
    return w < 0.5f ? a + w * (b - a) : b - (b - a) * (1.f - w);
}

__global__ void grad_sqnorm_kernel(const float* __restrict__ g, int64_t n, double* __restrict__ acc) {
    double s = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const double v = g[i];
        s += v * v;
    }
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    __shared__ double part[16];
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int i = 0; i < (int)(blockDim.x >> 6); i++) t += part[i];
        atomicAdd(acc, t);
    }
}

// torch.nn.utils.clip_grad_norm_: coef = max_norm / (norm + 1e-6), clamped to <= 1
// (NaN propagates, as torch.clamp does).  The accumulator is reset to zero for
// the next step, so a captured step needs no memset node.
__global__ void grad_clip_coef_kernel(double* __restrict__ acc, float max_norm, float* __restrict__ coef) {
    const float norm = (float)sqrt(*acc);
    const float c = max_norm / (norm + 1e-6f);
    *coef = c > 1.f ? 1.f : c;
    *acc = 0.0;
}

__global__ void adam_kernel(float* __restrict__ p, float* __restrict__ g, float* __restrict__ m,
                            float* __restrict__ v, int64_t n, int ncol, const float* __restrict__ col_lo,
                            const float* __restrict__ col_hi, double lr, const float* __restrict__ lr_dev,
                            double step, const float* __restrict__ step_dev, double beta1, double beta2, float eps,
                            float max_val, const float* __restrict__ coef, float lo_all, float hi_all) {
    __shared__ float slo[OPT_MAX_COLS], shi[OPT_MAX_COLS];
    __shared__ float sstep[2];
    if (col_lo) {
        for (int i = threadIdx.x; i < ncol; i += blockDim.x) {
            slo[i] = col_lo[i];
            shi[i] = col_hi[i];
        }
    }
    if (threadIdx.x == 0) {
        // bias corrections in double, as torch's Adam computes them in Python; the
        // step count and the learning rate may live on the device (graph replays)
        const double t = step_dev ? (double)*step_dev : step;
        const double rate = lr_dev ? (double)*lr_dev : lr;
        sstep[0] = (float)(-rate / (1.0 - pow(beta1, t)));
        sstep[1] = (float)sqrt(1.0 - pow(beta2, t));
    }
    __syncthreads();
    // scalars reach the elementwise math as float, as torch's tensor-scalar ops do
    const float step_size = sstep[0], bc2_sqrt = sstep[1];
    const float w1 = (float)(1.0 - beta1), b2 = (float)beta2, omb2 = (float)(1.0 - beta2);
    const float cf = coef ? *coef : 1.f;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        float gi = g[i];
        if (coef) gi *= cf;
        if (max_val > 0.f) gi = torch_clamp(gi, -max_val, max_val);
        gi = nan_to_num(gi);
        g[i] = gi;
        const float mi = torch_lerp(m[i], gi, w1);
        const float vi = v[i] * b2 + omb2 * gi * gi;
        m[i] = mi;
        v[i] = vi;
        const float denom = sqrtf(vi) / bc2_sqrt + eps;
        float pi = p[i] + step_size * (mi / denom);
        if (col_lo) {
            const int c = (int)(i % ncol);
            pi = torch_clamp(pi, slo[c], shi[c]);
        }
        pi = torch_clamp(pi, lo_all, hi_all);
        p[i] = pi;
    }
}

static int opt_blocks(int64_t n) {
    int64_t b = (n + 255) / 256;
    if (b > 4096) b = 4096;
    return (int)(b < 1 ? 1 : b);
}

}  // namespace n3ds

using namespace n3ds;

N3DS_EXPORT int n3ds_grad_sqnorm_accumulate(const float* grads, int64_t n, double* acc, void* stream) {
    N3DS_REQUIRE(n >= 0 && acc, "bad arguments");
    if (n == 0) return N3DS_OK;
    hipLaunchKernelGGL(grad_sqnorm_kernel, dim3(opt_blocks(n)), dim3(256), 0, (hipStream_t)stream, grads, n, acc);
    N3DS_LAUNCH_CHECK();
    return N3DS_OK;
}

N3DS_EXPORT int n3ds_grad_clip_coef(double* acc, float max_norm, float* coef, void* stream) {
    N3DS_REQUIRE(acc && coef && max_norm > 0.f, "bad arguments");
    hipLaunchKernelGGL(grad_clip_coef_kernel, dim3(1), dim3(1), 0, (hipStream_t)stream, acc, max_norm, coef);
    N3DS_LAUNCH_CHECK();
    return N3DS_OK;
}

N3DS_EXPORT int n3ds_adam_step(float* params, float* grads, float* exp_avg, float* exp_avg_sq, int64_t n_rows,
                               int n_cols, const float* col_lo, const float* col_hi, double lr, const float* lr_dev,
                               double beta1, double beta2, double eps, int64_t step, const float* step_dev,
                               float grad_max_val, const float* grad_coef, float lo_all, float hi_all, void* stream) {
    N3DS_REQUIRE(n_rows >= 0 && n_cols >= 1 && n_cols <= OPT_MAX_COLS, "bad sizes rows=%lld cols=%d",
                 (long long)n_rows, n_cols);
    N3DS_REQUIRE(step_dev || step >= 1, "step counts from 1 (got %lld)", (long long)step);
    N3DS_REQUIRE((col_lo == nullptr) == (col_hi == nullptr), "col_lo and col_hi go together");
    N3DS_REQUIRE(beta1 >= 0.0 && beta1 < 1.0 && beta2 >= 0.0 && beta2 < 1.0, "betas must lie in [0, 1)");
    const int64_t n = n_rows * n_cols;
    if (n == 0) return N3DS_OK;
    hipLaunchKernelGGL(adam_kernel, dim3(opt_blocks(n)), dim3(256), 0, (hipStream_t)stream, params, grads, exp_avg,
                       exp_avg_sq, n, n_cols, col_lo, col_hi, lr, lr_dev, (double)step, step_dev, beta1, beta2,
                       (float)eps, grad_max_val, grad_coef, lo_all, hi_all);
    N3DS_LAUNCH_CHECK();
    return N3DS_OK;
}
