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

// ---------------------------------------------------------------------------
// Level-ordered forward / atomic adjoint: a block owns GT consecutive points,
// one lane per point, and walks the levels in order, so the blocks resident at
// one time gather from one or two level tables (an XCD's 4 MB L2 holds one fine
// level in fp32) instead of all of them.  The [GT, L*C] slab of the output (or
// of the upstream gradient) passes through LDS so that global memory sees
// whole rows; LDS rows are padded to L*C + 1 floats against bank conflicts.
// ---------------------------------------------------------------------------
constexpr int GT = 256;
constexpr int GT_MAX_LC = 63;  // LDS slab of GT * (63 + 1) floats = 64 KB

// Forward: the block's [GT, L*C] slab of outputs is assembled in LDS and written
// as whole rows.  The LDS slab (33 KB at 16 levels x 2 features) also keeps the
// number of resident blocks, and with it the number of levels gathered from at
// one time, low: measured faster than keeping 64-byte runs of the row in
// registers at twice the occupancy (0.73 vs 0.82 ms at 2^20 points, fp32).
template <typename T, int D, int C>
__device__ __forceinline__ void grid_level_acc(const float* in, int level, const int* __restrict__ offsets,
                                               const LevelTab& lt, const T* __restrict__ grid, int gridtype,
                                               int align, int interp, float* acc) {
    float pos[D], deriv[D];
    uint32_t pg[D];
    const LevelCoord lc = level_coords<D>(in, level, offsets, lt, align != 0, interp, pos, deriv, pg);
#pragma unroll
    for (int c = 0; c < C; c++) acc[c] = 0.f;
    if (lc.oob) return;
    const T* g = grid + (int64_t)offsets[level] * C;
#pragma unroll
    for (int corner = 0; corner < (1 << D); corner++) {
        float w = 1.f;
        uint32_t loc[D];
#pragma unroll
        for (int d = 0; d < D; d++) {
            if (corner & (1 << d)) { w *= pos[d]; loc[d] = pg[d] + 1; }
            else { w *= 1.f - pos[d]; loc[d] = pg[d]; }
        }
        const uint32_t idx = grid_index<D>(gridtype, align != 0, lc.hs, lc.res, loc) * C;
#pragma unroll
        for (int c = 0; c < C; c++) acc[c] += w * to_f<T>(g[idx + c]);
    }
}

template <typename T, int D, int C>
__global__ __launch_bounds__(GT) void grid_fwd_tile_kernel(const float* __restrict__ inputs, const T* __restrict__ grid,
                                                           const int* __restrict__ offsets, T* __restrict__ outputs,
                                                           int64_t B, int L, LevelTab lt, int gridtype, int align,
                                                           int interp) {
    extern __shared__ float slab[];
    const int LC = L * C, row = LC + 1;
    const int64_t b0 = (int64_t)blockIdx.x * GT;
    const int nb = (int)min((int64_t)GT, B - b0);
    const int t = threadIdx.x;
    if (t < nb) {
        float in[D];
#pragma unroll
        for (int d = 0; d < D; d++) in[d] = inputs[(b0 + t) * D + d];
        // two levels per iteration: 2 x 2^D independent gathers in flight
#pragma unroll 1
        for (int level = 0; level < L; level += 2) {
            float a0[C], a1[C];
            grid_level_acc<T, D, C>(in, level, offsets, lt, grid, gridtype, align, interp, a0);
            if (level + 1 < L) grid_level_acc<T, D, C>(in, level + 1, offsets, lt, grid, gridtype, align, interp, a1);
#pragma unroll
            for (int c = 0; c < C; c++) slab[t * row + level * C + c] = a0[c];
            if (level + 1 < L) {
#pragma unroll
                for (int c = 0; c < C; c++) slab[t * row + (level + 1) * C + c] = a1[c];
            }
        }
    }
    __syncthreads();
    T* out = outputs + b0 * LC;
    for (int i = t; i < nb * LC; i += GT) {
        const int p = i / LC;
        out[i] = from_f<T>(slab[p * row + (i - p * LC)]);
    }
}

template <typename T, int D, int C>
__global__ __launch_bounds__(GT) void grid_bwd_tile_kernel(const T* __restrict__ grad, const float* __restrict__ inputs,
                                                           const int* __restrict__ offsets, float* __restrict__ ggrid,
                                                           int64_t B, int L, LevelTab lt, int gridtype, int align,
                                                           int interp) {
    extern __shared__ float slab[];
    const int LC = L * C, row = LC + 1;
    const int64_t b0 = (int64_t)blockIdx.x * GT;
    const int nb = (int)min((int64_t)GT, B - b0);
    const int t = threadIdx.x;
    const T* gsrc = grad + b0 * LC;
    for (int i = t; i < nb * LC; i += GT) {
        const int p = i / LC;
        slab[p * row + (i - p * LC)] = to_f<T>(gsrc[i]);
    }
    float in[D];
    if (t < nb) {
#pragma unroll
        for (int d = 0; d < D; d++) in[d] = inputs[(b0 + t) * D + d];
    }
    __syncthreads();
    if (t >= nb) return;
#pragma unroll 1
    for (int level = 0; level < L; level++) {
        float pos[D], deriv[D];
        uint32_t pg[D];
        const LevelCoord lc = level_coords<D>(in, level, offsets, lt, align != 0, interp, pos, deriv, pg);
        if (lc.oob) continue;
        float gcur[C];
#pragma unroll
        for (int c = 0; c < C; c++) gcur[c] = slab[t * row + level * C + c];
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
            const uint32_t idx = grid_index<D>(gridtype, align != 0, lc.hs, lc.res, loc) * C;
#pragma unroll
            for (int c = 0; c < C; c++) atomicAdd(gg + idx + c, w * gcur[c]);
        }
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

__global__ void grid_fill_int_kernel(int* __restrict__ p, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
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

// ---------------------------------------------------------------------------
// Bucketed gradient scatter (no float atomics to global memory on the hot path).
// Device-scope float atomics run at a fraction of HBM speed, so the adjoint
// groups its (row, value) contributions by bucket of RB consecutive table rows
// and sums each bucket in LDS:
//   1. count:   per (level, point tile) block, an LDS histogram of the buckets
//               its contributions fall into, added to the global bucket counts;
//   2. scan:    bucket starts and, per bucket, the number of SEG-record segments;
//   3. scatter: the same contributions again, written as (row, C values) records
//               into per-bucket ranges reserved with one atomic per (block, bucket);
//   4. reduce:  one block per bucket segment accumulates its records into an
//               LDS copy of the bucket's rows and adds it to the gradient table
//               (plain read-modify-write when the bucket has one segment).
// ---------------------------------------------------------------------------
constexpr int GB_THREADS = 256;
constexpr int GB_PPT = 8;           // points per thread in the count / scatter passes
constexpr int GB_HBINS = 4096;      // LDS histogram bins per block (buckets one level spans)
constexpr int64_t GB_SEG = 262144;  // records per reduce block

__host__ __device__ constexpr int gb_rows(int C) { return 16384 / C; }  // RB: 64 KB of fp32 rows in LDS

struct GridBucketWs {
    int* count;      // [NB]
    int* start;      // [NB + 1]
    int* cursor;     // [NB]
    int* segstart;   // [NB + 1]
    uint32_t* rows;  // [records]
    float* vals;     // [records * C]
};

static int64_t gb_num_buckets(int64_t n_rows, int C) { return (n_rows + gb_rows(C) - 1) / gb_rows(C); }

static int64_t gb_workspace_bytes(int64_t B, int D, int C, int L, int64_t n_rows) {
    const int64_t NB = gb_num_buckets(n_rows, C);
    const int64_t records = B * L * (1ll << D);
    int64_t head = (4 * NB + 2) * 4;
    head = (head + 255) / 256 * 256;
    return head + records * 4 + records * C * 4;
}

static GridBucketWs gb_carve(void* ws, int64_t B, int D, int C, int L, int64_t n_rows) {
    const int64_t NB = gb_num_buckets(n_rows, C);
    const int64_t records = B * L * (1ll << D);
    GridBucketWs w;
    int* ip = (int*)ws;
    w.count = ip;
    w.start = ip + NB;
    w.cursor = ip + 2 * NB + 1;
    w.segstart = ip + 3 * NB + 1;
    int64_t head = (4 * NB + 2) * 4;
    head = (head + 255) / 256 * 256;
    w.rows = (uint32_t*)((char*)ws + head);
    w.vals = (float*)((char*)ws + head + records * 4);
    return w;
}

// PASS 0: count, PASS 1: scatter.  blockIdx.y = level, blockIdx.x = tile of GB_THREADS * GB_PPT points.
template <typename T, int D, int C, int PASS>
__global__ __launch_bounds__(GB_THREADS) void grid_bucket_kernel(const T* __restrict__ grad,
                                                                 const float* __restrict__ inputs,
                                                                 const int* __restrict__ offsets, int64_t B, int L,
                                                                 LevelTab lt, int gridtype, int align, int interp,
                                                                 GridBucketWs ws) {
    constexpr int RB = gb_rows(C);
    __shared__ int hist[GB_HBINS];
    __shared__ int base[GB_HBINS];
    const int level = blockIdx.y;
    const int row0 = offsets[level], row1 = offsets[level + 1];
    const int b_lo = row0 / RB;
    const int nbins = row1 > row0 ? (row1 - 1) / RB - b_lo + 1 : 0;
    const bool local = nbins <= GB_HBINS;
    for (int i = threadIdx.x; i < GB_HBINS; i += GB_THREADS) hist[i] = 0;
    __syncthreads();
    const int64_t p_begin = (int64_t)blockIdx.x * GB_THREADS * GB_PPT;
    // pass 1 needs the histogram before it can place records: count first (both passes)
#pragma unroll 1
    for (int it = 0; it < GB_PPT; it++) {
        const int64_t b = p_begin + (int64_t)it * GB_THREADS + threadIdx.x;
        if (b >= B) break;
        float pos[D], deriv[D];
        uint32_t pg[D];
        LevelCoord lc = level_coords<D>(inputs + b * D, level, offsets, lt, align != 0, interp, pos, deriv, pg);
        if (lc.oob) continue;
#pragma unroll
        for (int corner = 0; corner < (1 << D); corner++) {
            uint32_t loc[D];
#pragma unroll
            for (int d = 0; d < D; d++) loc[d] = (corner & (1 << d)) ? pg[d] + 1 : pg[d];
            const int row = row0 + (int)grid_index<D>(gridtype, align != 0, lc.hs, lc.res, loc);
            const int bk = row / RB;
            if (local) atomicAdd(&hist[bk - b_lo], 1);
            else if (PASS == 0) atomicAdd(&ws.count[bk], 1);
        }
    }
    __syncthreads();
    if (PASS == 0) {
        if (local)
            for (int i = threadIdx.x; i < nbins; i += GB_THREADS)
                if (hist[i]) atomicAdd(&ws.count[b_lo + i], hist[i]);
        return;
    }
    // pass 1: reserve this block's range in each bucket, then write the records
    if (local) {
        for (int i = threadIdx.x; i < nbins; i += GB_THREADS) {
            base[i] = hist[i] ? atomicAdd(&ws.cursor[b_lo + i], hist[i]) : 0;
            hist[i] = 0;
        }
    }
    __syncthreads();
#pragma unroll 1
    for (int it = 0; it < GB_PPT; it++) {
        const int64_t b = p_begin + (int64_t)it * GB_THREADS + threadIdx.x;
        if (b >= B) break;
        float pos[D], deriv[D];
        uint32_t pg[D];
        LevelCoord lc = level_coords<D>(inputs + b * D, level, offsets, lt, align != 0, interp, pos, deriv, pg);
        if (lc.oob) continue;
        float gcur[C];
#pragma unroll
        for (int c = 0; c < C; c++) gcur[c] = to_f<T>(grad[(b * L + level) * C + c]);
#pragma unroll
        for (int corner = 0; corner < (1 << D); corner++) {
            float wgt = 1.f;
            uint32_t loc[D];
#pragma unroll
            for (int d = 0; d < D; d++) {
                if (corner & (1 << d)) { wgt *= pos[d]; loc[d] = pg[d] + 1; }
                else { wgt *= 1.f - pos[d]; loc[d] = pg[d]; }
            }
            const int row = row0 + (int)grid_index<D>(gridtype, align != 0, lc.hs, lc.res, loc);
            const int bk = row / RB;
            const int64_t slot = local ? (int64_t)base[bk - b_lo] + atomicAdd(&hist[bk - b_lo], 1)
                                       : (int64_t)atomicAdd(&ws.cursor[bk], 1);
            ws.rows[slot] = (uint32_t)row;
#pragma unroll
            for (int c = 0; c < C; c++) ws.vals[slot * C + c] = wgt * gcur[c];
        }
    }
}

// bucket starts, reservation cursors and per-bucket segment offsets (one block)
__global__ __launch_bounds__(1024) void grid_bucket_scan_kernel(GridBucketWs ws, int64_t NB) {
    __shared__ int sa[1024], sb[1024];
    __shared__ int carry_a, carry_b;
    if (threadIdx.x == 0) { carry_a = 0; carry_b = 0; }
    __syncthreads();
    for (int64_t t0 = 0; t0 < NB; t0 += 1024) {
        const int64_t i = t0 + threadIdx.x;
        const int cnt = i < NB ? ws.count[i] : 0;
        const int nseg = (int)((cnt + GB_SEG - 1) / GB_SEG);
        sa[threadIdx.x] = cnt;
        sb[threadIdx.x] = nseg;
        __syncthreads();
        for (int off = 1; off < 1024; off <<= 1) {
            int va = threadIdx.x >= off ? sa[threadIdx.x - off] : 0;
            int vb = threadIdx.x >= off ? sb[threadIdx.x - off] : 0;
            __syncthreads();
            sa[threadIdx.x] += va;
            sb[threadIdx.x] += vb;
            __syncthreads();
        }
        if (i < NB) {
            const int ea = carry_a + sa[threadIdx.x] - cnt;
            ws.start[i] = ea;
            ws.cursor[i] = ea;
            ws.segstart[i] = carry_b + sb[threadIdx.x] - nseg;
        }
        __syncthreads();
        if (threadIdx.x == 1023) {
            carry_a += sa[1023];
            carry_b += sb[1023];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        ws.start[NB] = carry_a;
        ws.segstart[NB] = carry_b;
    }
}

template <int C>
__global__ __launch_bounds__(GB_THREADS) void grid_bucket_reduce_kernel(GridBucketWs ws, int64_t NB, int64_t n_rows,
                                                                        float* __restrict__ ggrid) {
    constexpr int RB = gb_rows(C);
    __shared__ float acc[RB * C];
    const int64_t j = blockIdx.x;
    if (j >= ws.segstart[NB]) return;
    // bucket of segment j: last bucket with segstart <= j
    int64_t lo = 0, hi = NB - 1;
    while (lo < hi) {
        const int64_t mid = (lo + hi + 1) / 2;
        if (ws.segstart[mid] <= j) lo = mid;
        else hi = mid - 1;
    }
    const int64_t bk = lo;
    const int nseg = ws.segstart[bk + 1] - ws.segstart[bk];
    const int64_t r0 = ws.start[bk] + (j - ws.segstart[bk]) * GB_SEG;
    const int64_t r1 = min(r0 + GB_SEG, (int64_t)ws.start[bk + 1]);
    for (int i = threadIdx.x; i < RB * C; i += GB_THREADS) acc[i] = 0.f;
    __syncthreads();
    const int row_base = (int)(bk * RB);
    for (int64_t r = r0 + threadIdx.x; r < r1; r += GB_THREADS) {
        const int lr = (int)ws.rows[r] - row_base;
#pragma unroll
        for (int c = 0; c < C; c++) atomicAdd(&acc[lr * C + c], ws.vals[r * C + c]);
    }
    __syncthreads();
    const int64_t nr = min((int64_t)RB, n_rows - row_base);
    float* g = ggrid + (int64_t)row_base * C;
    for (int64_t i = threadIdx.x; i < nr * C; i += GB_THREADS) {
        const float v = acc[i];
        if (v == 0.f) continue;
        if (nseg == 1) g[i] += v;
        else atomicAdd(&g[i], v);
    }
}

// ---------------------------------------------------------------------------
// Fused proposal density field: the reference's PropMLP forward
// (source/models.py:286-523 with disable_rgb: _warp_coords, GridEncoder with
// bound 1, density_layer = Linear(F, 64) -> ReLU -> Linear(64, 1), then
// softplus(raw + density_bias)), F = L * C grid features, in one kernel; the
// adjoint recomputes features and hidden units per point instead of storing
// the [N, 64] activations, writes dL/dfeatures for the grid scatter and
// reduces the weight gradients in LDS.
// ---------------------------------------------------------------------------
constexpr int PF_HID = 64;
constexpr int PF_MAXF = 32;
constexpr int PF_FWD_THREADS = 256;
constexpr int PF_BWD_THREADS = 128;
constexpr int PF_CH = 16;  // hidden units per weight-gradient chunk

struct PropArgs {
    const float* coords;  // [N, 3] sample positions
    int64_t N;
    int warp;             // 0: none (coords in [-1, 1]), 1: contract then / 2, 2: divide by half the box side
    float inv_half_box;
    const void* emb;
    const int* offsets;
    int L;
    LevelTab lt;
    int gridtype, align, interp;
    const float* W1;  // [64, F]
    const float* b1;  // [64]
    const float* W2;  // [1, 64]
    const float* b2;  // [1]
    float dbias;
};

__device__ __forceinline__ void wave_fence() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// models.py:19-33 and coord.contract (utils/coord.py:5-11), then grid.py:158
// (inputs + bound) / (2 bound) with bound 1
__device__ __forceinline__ void prop_warp(const PropArgs& P, int64_t b, float* cw, float* x01) {
    float x[3] = {P.coords[b * 3], P.coords[b * 3 + 1], P.coords[b * 3 + 2]};
    if (P.warp == 1) {
        const float eps = 1.1920929e-07f;
        const float m2 = fmaxf(x[0] * x[0] + x[1] * x[1] + x[2] * x[2], eps);
        const float k = m2 <= 1.f ? 1.f : (2.f * sqrtf(m2) - 1.f) / m2;
#pragma unroll
        for (int d = 0; d < 3; d++) x[d] = (m2 <= 1.f ? x[d] : k * x[d]) / 2.f;
    } else if (P.warp == 2) {
#pragma unroll
        for (int d = 0; d < 3; d++) x[d] = x[d] * P.inv_half_box;
    }
#pragma unroll
    for (int d = 0; d < 3; d++) {
        cw[d] = x[d];
        x01[d] = (x[d] + 1.f) / 2.f;
    }
}

template <typename T, int C, int FM>
__device__ __forceinline__ void prop_features(const PropArgs& P, const float* x01, float* f) {
#pragma unroll
    for (int l = 0; l < FM / C; l++) {
        float a[C];
#pragma unroll
        for (int c = 0; c < C; c++) a[c] = 0.f;
        if (l < P.L) grid_level_acc<T, 3, C>(x01, l, P.offsets, P.lt, (const T*)P.emb, P.gridtype, P.align, P.interp, a);
#pragma unroll
        for (int c = 0; c < C; c++) f[l * C + c] = a[c];
    }
}

// torch.nn.functional.softplus (beta 1, threshold 20) and its derivative
__device__ __forceinline__ float softplus20(float x) { return x > 20.f ? x : log1pf(expf(x)); }
__device__ __forceinline__ float dsoftplus20(float x) { return x > 20.f ? 1.f : 1.f / (1.f + expf(-x)); }

template <int FM>
__device__ __forceinline__ float prop_raw(const float* sW1, const float* sb1, const float* sW2, float b2, int F,
                                          const float* f) {
    float raw = b2;
#pragma unroll 4
    for (int j = 0; j < PF_HID; j++) {
        float a = sb1[j];
#pragma unroll
        for (int i = 0; i < FM; i++)
            if (i < F) a += sW1[j * F + i] * f[i];
        raw += sW2[j] * fmaxf(a, 0.f);
    }
    return raw;
}

// FM: per-lane feature capacity (16 or 32 >= F)
// FX > 0: the feature count is FX (= FM) at compile time
template <typename T, int C, int FM, int FX = 0>
__global__ __launch_bounds__(PF_FWD_THREADS) void prop_fwd_kernel(PropArgs P, float* __restrict__ density,
                                                                  float* __restrict__ coord_out,
                                                                  float* __restrict__ feat_out,
                                                                  float* __restrict__ x01_out) {
    __shared__ float sW1[PF_HID * PF_MAXF], sb1[PF_HID], sW2[PF_HID];
    const int F = FX > 0 ? FX : P.L * C;
    for (int i = threadIdx.x; i < PF_HID * F; i += PF_FWD_THREADS) sW1[i] = P.W1[i];
    for (int i = threadIdx.x; i < PF_HID; i += PF_FWD_THREADS) {
        sb1[i] = P.b1[i];
        sW2[i] = P.W2[i];
    }
    __syncthreads();
    const float b2 = P.b2[0];
    for (int64_t b = (int64_t)blockIdx.x * PF_FWD_THREADS + threadIdx.x; b < P.N;
         b += (int64_t)gridDim.x * PF_FWD_THREADS) {
        float cw[3], x01[3], f[FM];
        prop_warp(P, b, cw, x01);
        prop_features<T, C, FM>(P, x01, f);
        const float raw = prop_raw<FM>(sW1, sb1, sW2, b2, F, f);
        density[b] = softplus20(raw + P.dbias);
        if (coord_out) {
#pragma unroll
            for (int d = 0; d < 3; d++) coord_out[b * 3 + d] = cw[d];
        }
        if (feat_out) {  // kept for the adjoint, which then gathers nothing
#pragma unroll
            for (int i = 0; i < FM; i++)
                if (i < F) feat_out[b * F + i] = f[i];
#pragma unroll
            for (int d = 0; d < 3; d++) x01_out[b * 3 + d] = x01[d];
        }
    }
}

template <typename T, int C, int FM, int FX = 0>
__global__ __launch_bounds__(PF_BWD_THREADS) void prop_bwd_kernel(PropArgs P, const float* __restrict__ gdens,
                                                                  const float* __restrict__ feat_in,
                                                                  float* __restrict__ gfeat, float* __restrict__ x01_out,
                                                                  float* __restrict__ gW1, float* __restrict__ gb1,
                                                                  float* __restrict__ gW2, float* __restrict__ gb2) {
    constexpr int NW = PF_BWD_THREADS / 64;
    __shared__ float sW1[PF_HID * PF_MAXF], sb1[PF_HID], sW2[PF_HID];
    __shared__ float aW1[PF_HID * PF_MAXF], ab1[PF_HID], aW2[PF_HID], ab2[1];  // block accumulators
    __shared__ float sf[NW][64][FM + 1];                                        // features, point-major
    __shared__ float sdh[NW][64][PF_CH + 1], sh[NW][64][PF_CH + 1];            // one chunk of hidden units
    __shared__ float sg[NW][64];
    const int F = FX > 0 ? FX : P.L * C;
    const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // uniform per wavefront
    for (int i = threadIdx.x; i < PF_HID * F; i += PF_BWD_THREADS) {
        sW1[i] = P.W1[i];
        aW1[i] = 0.f;
    }
    for (int i = threadIdx.x; i < PF_HID; i += PF_BWD_THREADS) {
        sb1[i] = P.b1[i];
        sW2[i] = P.W2[i];
        ab1[i] = 0.f;
        aW2[i] = 0.f;
    }
    if (threadIdx.x == 0) ab2[0] = 0.f;
    __syncthreads();
    const float b2 = P.b2[0];
    // every lane runs the same number of tiles (points past N contribute zeros)
    const int64_t ntile = (P.N + PF_BWD_THREADS - 1) / PF_BWD_THREADS;
    float gb2_lane = 0.f;
    for (int64_t tile = blockIdx.x; tile < ntile; tile += gridDim.x) {
        const int64_t b = tile * PF_BWD_THREADS + threadIdx.x;
        const bool live = b < P.N;
        float f[FM], df[FM], g = 0.f;
#pragma unroll
        for (int i = 0; i < FM; i++) { f[i] = 0.f; df[i] = 0.f; }
        if (live) {
            if (feat_in) {
#pragma unroll
                for (int i = 0; i < FM; i++)
                    if (i < F) f[i] = feat_in[b * F + i];
            } else {
                float cw[3], x01[3];
                prop_warp(P, b, cw, x01);
                prop_features<T, C, FM>(P, x01, f);
#pragma unroll
                for (int d = 0; d < 3; d++) x01_out[b * 3 + d] = x01[d];
            }
            const float raw = prop_raw<FM>(sW1, sb1, sW2, b2, F, f);
            g = gdens[b] * dsoftplus20(raw + P.dbias);
        }
        gb2_lane += g;
#pragma unroll
        for (int i = 0; i < FM; i++) sf[wave][lane][i] = f[i];
        sg[wave][lane] = g;
        for (int j0 = 0; j0 < PF_HID; j0 += PF_CH) {
            // per point: the chunk's hidden units, their adjoints and dL/df
#pragma unroll 4
            for (int jj = 0; jj < PF_CH; jj++) {
                const int j = j0 + jj;
                float a = sb1[j];
#pragma unroll
                for (int i = 0; i < FM; i++)
                    if (i < F) a += sW1[j * F + i] * f[i];
                const float dh = a > 0.f ? g * sW2[j] : 0.f;
#pragma unroll
                for (int i = 0; i < FM; i++)
                    if (i < F) df[i] += dh * sW1[j * F + i];
                sdh[wave][lane][jj] = dh;
                sh[wave][lane][jj] = fmaxf(a, 0.f);
            }
            wave_fence();
            // per hidden unit: sums over the wave's 64 points; lane = (unit jj, feature group ig)
            constexpr int KG = FM / 4;  // features per lane group
            const int jj = lane & (PF_CH - 1), ig = lane >> 4;  // 4 groups of KG features
            float w1[KG], bsum = 0.f, w2 = 0.f;
#pragma unroll
            for (int k = 0; k < KG; k++) w1[k] = 0.f;
            for (int q = 0; q < 64; q++) {
                const float dh = sdh[wave][q][jj];
#pragma unroll
                for (int k = 0; k < KG; k++) w1[k] += dh * sf[wave][q][ig * KG + k];
                if (ig == 0) {
                    bsum += dh;
                    w2 += sg[wave][q] * sh[wave][q][jj];
                }
            }
            const int j = j0 + jj;
#pragma unroll
            for (int k = 0; k < KG; k++)
                if (ig * KG + k < F) atomicAdd(&aW1[j * F + ig * KG + k], w1[k]);
            if (ig == 0) {
                atomicAdd(&ab1[j], bsum);
                atomicAdd(&aW2[j], w2);
            }
            wave_fence();
        }
        if (live) {
#pragma unroll
            for (int i = 0; i < FM; i++)
                if (i < F) gfeat[b * F + i] = df[i];
        }
    }
    // db2 = sum of g over all points
    float v = gb2_lane;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    if (lane == 0) atomicAdd(&ab2[0], v);
    __syncthreads();
    for (int i = threadIdx.x; i < PF_HID * F; i += PF_BWD_THREADS) atomicAdd(&gW1[i], aW1[i]);
    for (int i = threadIdx.x; i < PF_HID; i += PF_BWD_THREADS) {
        atomicAdd(&gb1[i], ab1[i]);
        atomicAdd(&gW2[i], aW2[i]);
    }
    if (threadIdx.x == 0) atomicAdd(&gb2[0], ab2[0]);
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
    const bool tiled = dy_dx == nullptr && L * C <= GT_MAX_LC;
    const dim3 tgrid((unsigned)((B + GT - 1) / GT));
    const size_t tlds = (size_t)GT * (L * C + 1) * sizeof(float);
#define LAUNCH_F(DD, CC)                                                                                         \
    if (tiled && dtype == 0)                                                                                     \
        hipLaunchKernelGGL((grid_fwd_tile_kernel<float, DD, CC>), tgrid, dim3(GT), tlds, s, inputs,              \
                           (const float*)embeddings, offsets, (float*)outputs, B, L, lt, gridtype, align_corners, \
                           interp);                                                                              \
    else if (tiled)                                                                                              \
        hipLaunchKernelGGL((grid_fwd_tile_kernel<__half, DD, CC>), tgrid, dim3(GT), tlds, s, inputs,             \
                           (const __half*)embeddings, offsets, (__half*)outputs, B, L, lt, gridtype,             \
                           align_corners, interp);                                                               \
    else if (dtype == 0)                                                                                         \
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

N3DS_EXPORT int64_t n3ds_grid_encode_backward_workspace_bytes(int64_t B, int D, int C, int L, int64_t n_rows) {
    if (B < 0 || D < 1 || D > 5 || C < 1 || C > 8 || L < 1 || n_rows < 0) return -1;
    return gb_workspace_bytes(B, D, C, L, n_rows);
}

// grad_embeddings_f32: fp32 accumulation table (zero-initialised by the caller);
// grad_embeddings_out: optional table of the embedding dtype that receives the
// rounded result (may alias grad_embeddings_f32 when dtype == float32).
// workspace (workspace_bytes >= n3ds_grid_encode_backward_workspace_bytes):
// bucketed scatter; NULL: one float atomic per contribution.
N3DS_EXPORT int n3ds_grid_encode_backward(const void* grad, const float* inputs, const int32_t* offsets,
                                          float* grad_embeddings_f32, void* grad_embeddings_out, int64_t n_rows,
                                          int64_t B, int D, int C, int L, float S, int H, const void* dy_dx,
                                          void* grad_inputs, int gridtype, int align_corners, int interp, int dtype,
                                          void* workspace, int64_t workspace_bytes, void* stream) {
    int st = grid_check(B, D, C, L, dtype);
    const LevelTab lt = make_level_tab(L, S, H);
    if (st) return st;
    hipStream_t s = (hipStream_t)stream;
    const bool bucketed = workspace != nullptr;
    if (bucketed)
        N3DS_REQUIRE(workspace_bytes >= gb_workspace_bytes(B, D, C, L, n_rows), "grid workspace too small");
    if (B > 0 && bucketed && n_rows > 0) {
        const GridBucketWs ws = gb_carve(workspace, B, D, C, L, n_rows);
        const int64_t NB = gb_num_buckets(n_rows, C);
        hipLaunchKernelGGL(grid_fill_int_kernel, dim3((unsigned)((NB + 255) / 256)), dim3(256), 0, s, ws.count, NB);
        N3DS_LAUNCH_CHECK();
        dim3 gpass((unsigned)((B + GB_THREADS * GB_PPT - 1) / (GB_THREADS * GB_PPT)), (unsigned)L);
        const int64_t max_segs = (B * L * (1ll << D) + GB_SEG - 1) / GB_SEG + NB;
#define LAUNCH_G(DD, CC)                                                                                          \
    if (dtype == 0) {                                                                                             \
        hipLaunchKernelGGL((grid_bucket_kernel<float, DD, CC, 0>), gpass, dim3(GB_THREADS), 0, s,                 \
                           (const float*)grad, inputs, offsets, B, L, lt, gridtype, align_corners, interp, ws);   \
        hipLaunchKernelGGL(grid_bucket_scan_kernel, dim3(1), dim3(1024), 0, s, ws, NB);                           \
        hipLaunchKernelGGL((grid_bucket_kernel<float, DD, CC, 1>), gpass, dim3(GB_THREADS), 0, s,                 \
                           (const float*)grad, inputs, offsets, B, L, lt, gridtype, align_corners, interp, ws);   \
    } else {                                                                                                      \
        hipLaunchKernelGGL((grid_bucket_kernel<__half, DD, CC, 0>), gpass, dim3(GB_THREADS), 0, s,                \
                           (const __half*)grad, inputs, offsets, B, L, lt, gridtype, align_corners, interp, ws);  \
        hipLaunchKernelGGL(grid_bucket_scan_kernel, dim3(1), dim3(1024), 0, s, ws, NB);                           \
        hipLaunchKernelGGL((grid_bucket_kernel<__half, DD, CC, 1>), gpass, dim3(GB_THREADS), 0, s,                \
                           (const __half*)grad, inputs, offsets, B, L, lt, gridtype, align_corners, interp, ws);  \
    }                                                                                                             \
    hipLaunchKernelGGL((grid_bucket_reduce_kernel<CC>), dim3((unsigned)max_segs), dim3(GB_THREADS), 0, s, ws, NB, \
                       n_rows, grad_embeddings_f32);
        N3DS_GRID_DISPATCH(LAUNCH_G)
#undef LAUNCH_G
        N3DS_LAUNCH_CHECK();
    }
    if (B > 0) {
        dim3 grid((unsigned)((B * L + 255) / 256));
        const bool tiled = L * C <= GT_MAX_LC;
        const dim3 tgrid((unsigned)((B + GT - 1) / GT));
        const size_t tlds = (size_t)GT * (L * C + 1) * sizeof(float);
#define LAUNCH_B(DD, CC)                                                                                          \
    if (!bucketed && tiled && dtype == 0)                                                                         \
        hipLaunchKernelGGL((grid_bwd_tile_kernel<float, DD, CC>), tgrid, dim3(GT), tlds, s, (const float*)grad,   \
                           inputs, offsets, grad_embeddings_f32, B, L, lt, gridtype, align_corners, interp);      \
    else if (!bucketed && tiled)                                                                                  \
        hipLaunchKernelGGL((grid_bwd_tile_kernel<__half, DD, CC>), tgrid, dim3(GT), tlds, s, (const __half*)grad, \
                           inputs, offsets, grad_embeddings_f32, B, L, lt, gridtype, align_corners, interp);      \
    else if (!bucketed && dtype == 0)                                                                             \
        hipLaunchKernelGGL((grid_bwd_kernel<float, DD, CC>), grid, dim3(256), 0, s, (const float*)grad, inputs,   \
                           offsets, grad_embeddings_f32, B, L, lt, gridtype, align_corners, interp);   \
    else if (!bucketed)                                                                                           \
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

// ---------------------------------------------------------------------------
// fused proposal density field entry points
// ---------------------------------------------------------------------------
static int prop_check(int64_t N, int C, int L, int dtype, int warp) {
    N3DS_REQUIRE(N >= 0, "bad point count");
    N3DS_REQUIRE(C == 1 || C == 2 || C == 4, "level_dim must be 1, 2 or 4 (got %d)", C);
    N3DS_REQUIRE(L >= 1 && L * C <= PF_MAXF, "num_levels * level_dim must be in [1, %d] (got %d x %d)", PF_MAXF, L, C);
    N3DS_REQUIRE(dtype == 0 || dtype == 1, "dtype must be 0 (float32) or 1 (float16)");
    N3DS_REQUIRE(warp >= 0 && warp <= 2, "warp must be 0 (none), 1 (contract) or 2 (box)");
    return N3DS_OK;
}

static int g_pf_cu = 0;

static int prop_blocks(int64_t work, int threads, int per_cu, int64_t* out) {
    if (g_pf_cu == 0) {
        int dev;
        N3DS_HIP_CHECK(hipGetDevice(&dev));
        N3DS_HIP_CHECK(hipDeviceGetAttribute(&g_pf_cu, hipDeviceAttributeMultiprocessorCount, dev));
    }
    int64_t nb = (work + threads - 1) / threads;
    const int64_t cap = (int64_t)g_pf_cu * per_cu;
    *out = nb < 1 ? 1 : (nb > cap ? cap : nb);
    return N3DS_OK;
}

static PropArgs prop_args(const float* coords, int64_t N, int warp, float box_size, const void* emb,
                          const int32_t* offsets, int L, float S, int H, int gridtype, int align, int interp,
                          const float* W1, const float* b1, const float* W2, const float* b2, float dbias) {
    PropArgs P;
    P.coords = coords; P.N = N; P.warp = warp; P.inv_half_box = box_size > 0.f ? 2.f / box_size : 1.f;
    P.emb = emb; P.offsets = offsets; P.L = L; P.lt = make_level_tab(L, S, H);
    P.gridtype = gridtype; P.align = align; P.interp = interp;
    P.W1 = W1; P.b1 = b1; P.W2 = W2; P.b2 = b2; P.dbias = dbias;
    return P;
}

N3DS_EXPORT int n3ds_prop_field_forward(const float* coords, int64_t n_points, int warp, float box_size,
                                        const void* embeddings, const int32_t* offsets, int num_levels,
                                        int level_dim, float S, int H, int gridtype, int align_corners, int interp,
                                        int dtype, const float* w1, const float* b1, const float* w2,
                                        const float* b2, float density_bias, float* density_out,
                                        float* coord_out, float* features_out, float* grid_inputs_out,
                                        void* stream) {
    N3DS_REQUIRE((features_out == nullptr) == (grid_inputs_out == nullptr),
                 "features_out and grid_inputs_out are kept together");
    int st = prop_check(n_points, level_dim, num_levels, dtype, warp);
    if (st) return st;
    if (n_points == 0) return N3DS_OK;
    const PropArgs P = prop_args(coords, n_points, warp, box_size, embeddings, offsets, num_levels, S, H, gridtype,
                                 align_corners, interp, w1, b1, w2, b2, density_bias);
    int64_t nb;
    st = prop_blocks(n_points, PF_FWD_THREADS, 8, &nb);
    if (st) return st;
    hipStream_t s = (hipStream_t)stream;
#define LAUNCH_PF(CC, FM, FX)                                                                                    \
    if (dtype == 0)                                                                                              \
        hipLaunchKernelGGL((prop_fwd_kernel<float, CC, FM, FX>), dim3((unsigned)nb), dim3(PF_FWD_THREADS), 0, s, \
                           P, density_out, coord_out, features_out, grid_inputs_out);                            \
    else                                                                                                         \
        hipLaunchKernelGGL((prop_fwd_kernel<__half, CC, FM, FX>), dim3((unsigned)nb), dim3(PF_FWD_THREADS), 0,   \
                           s, P, density_out, coord_out, features_out, grid_inputs_out);
    const bool small = num_levels * level_dim <= 16;
    if (level_dim == 1 && num_levels == 10) {
        LAUNCH_PF(1, 10, 10)
    } else if (level_dim == 2 && num_levels == 10) {
        LAUNCH_PF(2, 20, 20)
    } else if (level_dim == 1) {
        if (small) { LAUNCH_PF(1, 16, 0) } else { LAUNCH_PF(1, 32, 0) }
    } else if (level_dim == 2) {
        if (small) { LAUNCH_PF(2, 16, 0) } else { LAUNCH_PF(2, 32, 0) }
    } else {
        if (small) { LAUNCH_PF(4, 16, 0) } else { LAUNCH_PF(4, 32, 0) }
    }
#undef LAUNCH_PF
    N3DS_LAUNCH_CHECK();
    return N3DS_OK;
}

N3DS_EXPORT int n3ds_prop_field_backward(const float* coords, int64_t n_points, int warp, float box_size,
                                         const void* embeddings, const int32_t* offsets, int num_levels,
                                         int level_dim, float S, int H, int gridtype, int align_corners, int interp,
                                         int dtype, const float* w1, const float* b1, const float* w2,
                                         const float* b2, float density_bias, const float* grad_density,
                                         const float* features, float* grad_features, float* grid_inputs,
                                         float* grad_w1, float* grad_b1, float* grad_w2, float* grad_b2,
                                         void* stream) {
    int st = prop_check(n_points, level_dim, num_levels, dtype, warp);
    if (st) return st;
    if (n_points == 0) return N3DS_OK;
    const PropArgs P = prop_args(coords, n_points, warp, box_size, embeddings, offsets, num_levels, S, H, gridtype,
                                 align_corners, interp, w1, b1, w2, b2, density_bias);
    int64_t nb;
    st = prop_blocks(n_points, PF_BWD_THREADS, 3, &nb);  // 3 resident blocks per CU (43 KB of LDS each)
    if (st) return st;
    hipStream_t s = (hipStream_t)stream;
#define LAUNCH_PB(CC, FM, FX)                                                                                    \
    if (dtype == 0)                                                                                              \
        hipLaunchKernelGGL((prop_bwd_kernel<float, CC, FM, FX>), dim3((unsigned)nb), dim3(PF_BWD_THREADS), 0, s, \
                           P, grad_density, features, grad_features, grid_inputs, grad_w1, grad_b1, grad_w2,     \
                           grad_b2);                                                                             \
    else                                                                                                         \
        hipLaunchKernelGGL((prop_bwd_kernel<__half, CC, FM, FX>), dim3((unsigned)nb), dim3(PF_BWD_THREADS), 0,   \
                           s, P, grad_density, features, grad_features, grid_inputs, grad_w1, grad_b1, grad_w2,  \
                           grad_b2);
    const bool small = num_levels * level_dim <= 16;
    if (level_dim == 1 && num_levels == 10) {
        LAUNCH_PB(1, 12, 10)
    } else if (level_dim == 2 && num_levels == 10) {
        LAUNCH_PB(2, 20, 20)
    } else if (level_dim == 1) {
        if (small) { LAUNCH_PB(1, 16, 0) } else { LAUNCH_PB(1, 32, 0) }
    } else if (level_dim == 2) {
        if (small) { LAUNCH_PB(2, 16, 0) } else { LAUNCH_PB(2, 32, 0) }
    } else {
        if (small) { LAUNCH_PB(4, 16, 0) } else { LAUNCH_PB(4, 32, 0) }
    }
#undef LAUNCH_PB
    N3DS_LAUNCH_CHECK();
    return N3DS_OK;
}
