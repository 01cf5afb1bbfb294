// Multiresolution hash-grid encoder (drop-in for source/gridencoder/src/gridencoder.cu,
// bound in source/gridencoder/grid.py:54,84,194).
//
// Layout: one lane per (point, level) with the level fastest, so a wavefront
// writes a contiguous span of the [B, L*C] output (the reference writes
// [L, B, C] and permutes in Python, grid.py:57).  Table gathers read C
// contiguous features per corner (float2 / half2 for C = 2).  fp16 tables are
// interpolated in fp32 and rounded once on store; the adjoint accumulates into
// an fp32 gradient table (the reference accumulates fp16 with __half2 atomics).
// Index math is uint32 with wrap-around exactly as gridencoder.cu:50-84.
#include <hip/hip_fp16.h>

#include "n3ds_capi.h"

namespace n3ds {

__device__ __forceinline__ uint32_t grid_hash(const uint32_t* pg, int D) {
    const uint32_t primes[7] = {1u, 2654435761u, 805459861u, 3674653429u, 2097192037u, 1434869437u, 2165219737u};
    uint32_t r = 0;
    for (int i = 0; i < D; i++) r ^= pg[i] * primes[i];
    return r;
}

template <int D>
__device__ __forceinline__ uint32_t grid_index(int gridtype, bool align, uint32_t hs, uint32_t res, const uint32_t* pg) {
    uint32_t stride = 1, index = 0;
#pragma unroll
    for (int d = 0; d < D; d++) {
        if (stride <= hs) {
            index += pg[d] * stride;
            stride *= align ? res : (res + 1);
        }
    }
    if (gridtype == 0 && stride > hs) {
        const uint32_t primes[7] = {1u, 2654435761u, 805459861u, 3674653429u, 2097192037u, 1434869437u, 2165219737u};
        index = 0;
#pragma unroll
        for (int d = 0; d < D; d++) index ^= pg[d] * primes[d];
    }
    return index % hs;
}

template <typename T>
__device__ __forceinline__ float to_f(T v);
template <>
__device__ __forceinline__ float to_f<float>(float v) { return v; }
template <>
__device__ __forceinline__ float to_f<__half>(__half v) { return __half2float(v); }
template <typename T>
__device__ __forceinline__ T from_f(float v);
template <>
__device__ __forceinline__ float from_f<float>(float v) { return v; }
template <>
__device__ __forceinline__ __half from_f<__half>(float v) { return __float2half(v); }

struct LevelCoord {
    uint32_t hs, res;
    float scale;
    bool oob;
};

constexpr int MAX_LEVELS = 64;
// Per-level grid scale exp2(level*S)*H - 1 (gridencoder.cu:138), evaluated once
// on the host with float32 operands and a correctly rounded exp2, so integer
// powers of two are exact.
struct LevelTab {
    float scale[MAX_LEVELS];
};

static LevelTab make_level_tab(int L, float S, int H) {
    LevelTab t;
    for (int l = 0; l < MAX_LEVELS; l++) {
        float ls = (float)l * S;
        float e = (float)exp2((double)ls);
        t.scale[l] = l < L ? (e * (float)H) - 1.0f : 0.f;
    }
    return t;
}

template <int D>
__device__ __forceinline__ LevelCoord level_coords(const float* in, int level, const int* offsets, const LevelTab& lt,
                                                   bool align, int interp, float* pos, float* deriv, uint32_t* pg) {
    LevelCoord lc;
    lc.oob = false;
#pragma unroll
    for (int d = 0; d < D; d++)
        if (in[d] < 0.f || in[d] > 1.f) lc.oob = true;
    lc.hs = (uint32_t)(offsets[level + 1] - offsets[level]);
    if (lc.hs == 0) lc.oob = true;  // empty level: nothing to gather (avoids a modulo by zero)
    lc.scale = lt.scale[level];
    lc.res = (uint32_t)ceilf(lc.scale) + 1;
#pragma unroll
    for (int d = 0; d < D; d++) {
        // one rounding, as nvcc contracts the reference's x * scale + 0.5 (at the
        // finest levels the float32 ulp of pos is 1/16, so double rounding matters)
        float p = fmaf(in[d], lc.scale, align ? 0.0f : 0.5f);
        float f = floorf(p);
        pg[d] = (uint32_t)f;
        p -= f;
        if (interp == 1) {
            deriv[d] = 6.f * p * (1.f - p);
            p = p * p * (3.f - 2.f * p);
        } else {
            deriv[d] = 1.f;
        }
        pos[d] = p;
    }
    return lc;
}

template <typename T, int D, int C>
__global__ __launch_bounds__(256) void grid_fwd_kernel(const float* __restrict__ inputs, const T* __restrict__ grid,
                                                       const int* __restrict__ offsets, T* __restrict__ outputs,
                                                       int64_t B, int L, LevelTab lt, T* __restrict__ dy_dx,
                                                       int gridtype, int align, int interp) {
    const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (tid >= B * L) return;
    const int64_t b = tid / L;
    const int level = (int)(tid - b * L);
    const float* in = inputs + b * D;
    float pos[D], deriv[D];
    uint32_t pg[D];
    LevelCoord lc = level_coords<D>(in, level, offsets, lt, align != 0, interp, pos, deriv, pg);
    T* out = outputs + tid * C;
    if (lc.oob) {
#pragma unroll
        for (int c = 0; c < C; c++) out[c] = from_f<T>(0.f);
        if (dy_dx) {
            T* dd = dy_dx + b * (int64_t)L * D * C + (int64_t)level * D * C;
            for (int i = 0; i < D * C; i++) dd[i] = from_f<T>(0.f);
        }
        return;
    }
    const T* g = grid + (int64_t)offsets[level] * C;
    float acc[C];
#pragma unroll
    for (int c = 0; c < C; c++) acc[c] = 0.f;
#pragma unroll
    for (int corner = 0; corner < (1 << D); corner++) {
        float w = 1.f;
        uint32_t loc[D];
#pragma unroll
        for (int d = 0; d < D; d++) {
            if (corner & (1 << d)) { w *= pos[d]; loc[d] = pg[d] + 1; }
            else { w *= 1.f - pos[d]; loc[d] = pg[d]; }
        }
        uint32_t idx = grid_index<D>(gridtype, align != 0, lc.hs, lc.res, loc) * C;
#pragma unroll
        for (int c = 0; c < C; c++) acc[c] += w * to_f<T>(g[idx + c]);
    }
#pragma unroll
    for (int c = 0; c < C; c++) out[c] = from_f<T>(acc[c]);
    if (dy_dx) {
        T* dd = dy_dx + b * (int64_t)L * D * C + (int64_t)level * D * C;
#pragma unroll
        for (int gd = 0; gd < D; gd++) {
            float rg[C];
#pragma unroll
            for (int c = 0; c < C; c++) rg[c] = 0.f;
#pragma unroll
            for (int corner = 0; corner < (1 << (D - 1)); corner++) {
                float w = lc.scale;
                uint32_t loc[D];
#pragma unroll
                for (int nd = 0; nd < D - 1; nd++) {
                    int d = nd >= gd ? nd + 1 : nd;
                    if (corner & (1 << nd)) { w *= pos[d]; loc[d] = pg[d] + 1; }
                    else { w *= 1.f - pos[d]; loc[d] = pg[d]; }
                }
                loc[gd] = pg[gd];
                uint32_t il = grid_index<D>(gridtype, align != 0, lc.hs, lc.res, loc) * C;
                loc[gd] = pg[gd] + 1;
                uint32_t ir = grid_index<D>(gridtype, align != 0, lc.hs, lc.res, loc) * C;
#pragma unroll
                for (int c = 0; c < C; c++) rg[c] += w * (to_f<T>(g[ir + c]) - to_f<T>(g[il + c])) * deriv[gd];
            }
#pragma unroll
            for (int c = 0; c < C; c++) dd[gd * C + c] = from_f<T>(rg[c]);
        }
    }
}

template <typename T, int D, int C>
__global__ __launch_bounds__(256) void grid_bwd_kernel(const T* __restrict__ grad, const float* __restrict__ inputs,
                                                       const int* __restrict__ offsets, float* __restrict__ ggrid,
                                                       int64_t B, int L, LevelTab lt, int gridtype, int align,
                                                       int interp) {
    const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (tid >= B * L) return;
    const int64_t b = tid / L;
    const int level = (int)(tid - b * L);
    const float* in = inputs + b * D;
    float pos[D], deriv[D];
    uint32_t pg[D];
    LevelCoord lc = level_coords<D>(in, level, offsets, lt, align != 0, interp, pos, deriv, pg);
    if (lc.oob) return;
    float gcur[C];
#pragma unroll
    for (int c = 0; c < C; c++) gcur[c] = to_f<T>(grad[tid * C + c]);
    float* gg = ggrid + (int64_t)offsets[level] * C;
#pragma unroll
    for (int corner = 0; corner < (1 << D); corner++) {
        float w = 1.f;
        uint32_t loc[D];
#pragma unroll
        for (int d = 0; d < D; d++) {
            if (corner & (1 << d)) { w *= pos[d]; loc[d] = pg[d] + 1; }
            else { w *= 1.f - pos[d]; loc[d] = pg[d]; }
        }
        uint32_t idx = grid_index<D>(gridtype, align != 0, lc.hs, lc.res, loc) * C;
#pragma unroll
        for (int c = 0; c < C; c++) atomicAdd(gg + idx + c, w * gcur[c]);
    }
}

// dL/dx[b, d] = sum_{l, c} grad[b, l, c] * dy_dx[b, l, d, c]   (gridencoder.cu:343-369)
template <typename T, int D, int C>
__global__ void grid_input_bwd_kernel(const T* __restrict__ grad, const T* __restrict__ dy_dx, T* __restrict__ gin,
                                      int64_t B, int L) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= B * D) return;
    const int64_t b = t / D;
    const int d = (int)(t - b * D);
    float r = 0.f;
    for (int l = 0; l < L; l++)
#pragma unroll
        for (int c = 0; c < C; c++)
            r += to_f<T>(grad[(b * L + l) * C + c]) * to_f<T>(dy_dx[b * (int64_t)L * D * C + (int64_t)l * D * C + d * C + c]);
    gin[t] = from_f<T>(r);
}

template <typename T>
__global__ void f32_to_kernel(const float* __restrict__ src, T* __restrict__ dst, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = from_f<T>(src[i]);
}

// total-variation gradient added into grad (gridencoder.cu:506-610)
template <typename T, int D, int C>
__global__ void grid_tv_kernel(const float* __restrict__ inputs, const T* __restrict__ grid, float* __restrict__ grad,
                               const int* __restrict__ offsets, float weight, int64_t B, int L, LevelTab lt,
                               int gridtype, int align) {
    const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (tid >= B * L) return;
    const int64_t b = tid / L;
    const int level = (int)(tid - b * L);
    const float* in = inputs + b * D;
    float pos[D], deriv[D];
    uint32_t pg[D];
    LevelCoord lc = level_coords<D>(in, level, offsets, lt, align != 0, 0, pos, deriv, pg);
    if (lc.oob) return;
    const T* g = grid + (int64_t)offsets[level] * C;
    float* gr = grad + (int64_t)offsets[level] * C;
    uint32_t idx = grid_index<D>(gridtype, align != 0, lc.hs, lc.res, pg) * C;
    float res[C], idel[C];
#pragma unroll
    for (int c = 0; c < C; c++) { res[c] = 0.f; idel[c] = 0.f; }
    const float w = weight / (2 * D);
#pragma unroll
    for (int d = 0; d < D; d++) {
        uint32_t cur = pg[d];
        if (cur < lc.res) {
            pg[d] = cur + 1;
            uint32_t ir = grid_index<D>(gridtype, align != 0, lc.hs, lc.res, pg) * C;
#pragma unroll
            for (int c = 0; c < C; c++) {
                float v = to_f<T>(g[idx + c]) - to_f<T>(g[ir + c]);
                res[c] += v;
                idel[c] += v * v;
            }
        }
        if (cur > 0) {
            pg[d] = cur - 1;
            uint32_t il = grid_index<D>(gridtype, align != 0, lc.hs, lc.res, pg) * C;
#pragma unroll
            for (int c = 0; c < C; c++) {
                float v = to_f<T>(g[idx + c]) - to_f<T>(g[il + c]);
                res[c] += v;
                idel[c] += v * v;
            }
        }
        pg[d] = cur;
    }
#pragma unroll
    for (int c = 0; c < C; c++) atomicAdd(gr + idx + c, w * res[c] * rsqrtf(idel[c] + 1e-9f));
}

#define N3DS_GRID_DISPATCH(KERNEL_LAUNCH)                                                           \
    switch (D * 16 + C) {                                                                          \
        case 2 * 16 + 1: KERNEL_LAUNCH(2, 1); break;                                               \
        case 2 * 16 + 2: KERNEL_LAUNCH(2, 2); break;                                               \
        case 2 * 16 + 4: KERNEL_LAUNCH(2, 4); break;                                               \
        case 2 * 16 + 8: KERNEL_LAUNCH(2, 8); break;                                               \
        case 3 * 16 + 1: KERNEL_LAUNCH(3, 1); break;                                               \
        case 3 * 16 + 2: KERNEL_LAUNCH(3, 2); break;                                               \
        case 3 * 16 + 4: KERNEL_LAUNCH(3, 4); break;                                               \
        case 3 * 16 + 8: KERNEL_LAUNCH(3, 8); break;                                               \
        case 4 * 16 + 2: KERNEL_LAUNCH(4, 2); break;                                               \
        case 5 * 16 + 2: KERNEL_LAUNCH(5, 2); break;                                               \
        default: N3DS_REQUIRE(false, "unsupported input_dim %d / level_dim %d", D, C);             \
    }

static int grid_check(int64_t B, int D, int C, int L, int dtype) {
    N3DS_REQUIRE(B >= 0 && L > 0 && L <= MAX_LEVELS, "bad sizes B=%lld L=%d", (long long)B, L);
    N3DS_REQUIRE(D >= 2 && D <= 5, "input_dim must be in [2, 5] (got %d)", D);
    N3DS_REQUIRE(C == 1 || C == 2 || C == 4 || C == 8, "level_dim must be 1, 2, 4 or 8 (got %d)", C);
    N3DS_REQUIRE(dtype == 0 || dtype == 1, "dtype must be 0 (float32) or 1 (float16)");
    return N3DS_OK;
}

}  // namespace n3ds

using namespace n3ds;

N3DS_EXPORT int n3ds_grid_encode_forward(const float* inputs, const void* embeddings, const int32_t* offsets,
                                         void* outputs, int64_t B, int D, int C, int L, float S, int H, void* dy_dx,
                                         int gridtype, int align_corners, int interp, int dtype, void* stream) {
    int st = grid_check(B, D, C, L, dtype);
    const LevelTab lt = make_level_tab(L, S, H);
    if (st) return st;
    if (B == 0) return N3DS_OK;
    dim3 grid((unsigned)((B * L + 255) / 256));
    hipStream_t s = (hipStream_t)stream;
#define LAUNCH_F(DD, CC)                                                                                         \
    if (dtype == 0)                                                                                              \
        hipLaunchKernelGGL((grid_fwd_kernel<float, DD, CC>), grid, dim3(256), 0, s, inputs, (const float*)embeddings, \
                           offsets, (float*)outputs, B, L, lt, (float*)dy_dx, gridtype, align_corners, interp); \
    else                                                                                                         \
        hipLaunchKernelGGL((grid_fwd_kernel<__half, DD, CC>), grid, dim3(256), 0, s, inputs,                     \
                           (const __half*)embeddings, offsets, (__half*)outputs, B, L, lt,            \
                           (__half*)dy_dx, gridtype, align_corners, interp);
    N3DS_GRID_DISPATCH(LAUNCH_F)
#undef LAUNCH_F
    N3DS_LAUNCH_CHECK();
    return N3DS_OK;
}

// grad_embeddings_f32: fp32 accumulation table (zero-initialised by the caller);
// grad_embeddings_out: optional table of the embedding dtype that receives the
// rounded result (may alias grad_embeddings_f32 when dtype == float32).
N3DS_EXPORT int n3ds_grid_encode_backward(const void* grad, const float* inputs, const int32_t* offsets,
                                          float* grad_embeddings_f32, void* grad_embeddings_out, int64_t n_rows,
                                          int64_t B, int D, int C, int L, float S, int H, const void* dy_dx,
                                          void* grad_inputs, int gridtype, int align_corners, int interp, int dtype,
                                          void* stream) {
    int st = grid_check(B, D, C, L, dtype);
    const LevelTab lt = make_level_tab(L, S, H);
    if (st) return st;
    hipStream_t s = (hipStream_t)stream;
    if (B > 0) {
        dim3 grid((unsigned)((B * L + 255) / 256));
#define LAUNCH_B(DD, CC)                                                                                          \
    if (dtype == 0)                                                                                               \
        hipLaunchKernelGGL((grid_bwd_kernel<float, DD, CC>), grid, dim3(256), 0, s, (const float*)grad, inputs,   \
                           offsets, grad_embeddings_f32, B, L, lt, gridtype, align_corners, interp);   \
    else                                                                                                          \
        hipLaunchKernelGGL((grid_bwd_kernel<__half, DD, CC>), grid, dim3(256), 0, s, (const __half*)grad, inputs, \
                           offsets, grad_embeddings_f32, B, L, lt, gridtype, align_corners, interp);   \
    if (dy_dx && grad_inputs) {                                                                                   \
        dim3 g2((unsigned)((B * DD + 255) / 256));                                                                \
        if (dtype == 0)                                                                                           \
            hipLaunchKernelGGL((grid_input_bwd_kernel<float, DD, CC>), g2, dim3(256), 0, s, (const float*)grad,   \
                               (const float*)dy_dx, (float*)grad_inputs, B, L);                                   \
        else                                                                                                      \
            hipLaunchKernelGGL((grid_input_bwd_kernel<__half, DD, CC>), g2, dim3(256), 0, s, (const __half*)grad, \
                               (const __half*)dy_dx, (__half*)grad_inputs, B, L);                                 \
    }
        N3DS_GRID_DISPATCH(LAUNCH_B)
#undef LAUNCH_B
        N3DS_LAUNCH_CHECK();
    }
    if (grad_embeddings_out && dtype == 1 && n_rows > 0) {
        int64_t n = n_rows * C;
        hipLaunchKernelGGL(f32_to_kernel<__half>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s,
                           grad_embeddings_f32, (__half*)grad_embeddings_out, n);
        N3DS_LAUNCH_CHECK();
    }
    return N3DS_OK;
}

N3DS_EXPORT int n3ds_grid_grad_tv(const float* inputs, const void* embeddings, float* grad_f32, const int32_t* offsets,
                                  float weight, int64_t B, int D, int C, int L, float S, int H, int gridtype,
                                  int align_corners, int dtype, void* stream) {
    int st = grid_check(B, D, C, L, dtype);
    const LevelTab lt = make_level_tab(L, S, H);
    if (st) return st;
    if (B == 0) return N3DS_OK;
    dim3 grid((unsigned)((B * L + 255) / 256));
    hipStream_t s = (hipStream_t)stream;
#define LAUNCH_T(DD, CC)                                                                                         \
    if (dtype == 0)                                                                                              \
        hipLaunchKernelGGL((grid_tv_kernel<float, DD, CC>), grid, dim3(256), 0, s, inputs, (const float*)embeddings, \
                           grad_f32, offsets, weight, B, L, lt, gridtype, align_corners);             \
    else                                                                                                         \
        hipLaunchKernelGGL((grid_tv_kernel<__half, DD, CC>), grid, dim3(256), 0, s, inputs,                      \
                           (const __half*)embeddings, grad_f32, offsets, weight, B, L, lt, gridtype,  \
                           align_corners);
    N3DS_GRID_DISPATCH(LAUNCH_T)
#undef LAUNCH_T
    N3DS_LAUNCH_CHECK();
    return N3DS_OK;
}
