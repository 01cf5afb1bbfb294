// Image texture on stroke colours: drop-in replacement for the torch ops of the
// reference's ImageTexture (source/textures.py:36-82):
//   tint   = grid_sample(image, texcoords * 2 - 1, bilinear, zero padding,
//                        align_corners=True)                      (:51-61)
//   colour = (colour * tint_ratio + (1 - tint_ratio)) * tint        (:77-78)
//   alpha  = alpha * mean(tint)          when modulate_alpha        (:79-80)
//
// Layout: the image [C][H][W] (the reference's buffer) is packed once into
// "quads": for every cell origin (y0, x0), y0 in [-1, H-1], x0 in [-1, W-1],
// one entry holds the four corners nw (x0,y0), ne (x0+1,y0), sw (x0,y0+1),
// se (x0+1,y0+1) channel by channel, with zeros for corners off the image
// (zero padding).  An RGB entry is 12 floats padded to 64 B, a grey entry 4
// floats (16 B), so a bilinear tap reads one aligned segment of one cache
// line instead of four scattered texels; a 256x256 RGB texture packs to
// 4.2 MB, resident in L2 / MALL.  A stored zero times a finite weight adds
// +0, so the result equals skipping the off-image corners.  One thread per
// (sample, stroke) pair in a grid-stride loop; per pair the kernel streams
// 8 B of texcoords, 12 B of colour in and out, and 4 B of alpha in and out
// when modulating.
//
// Derivatives are those torch's autograd gives the same ops (the image is a
// buffer without gradient): d colour_in = g * tint * tint_ratio,
// d alpha_in = g_alpha * mean(tint), and d texcoords from the bilinear
// weights with out-of-range corners contributing zero on both sides.
#include <algorithm>

#include "n3ds_capi.h"
#include "n3ds_math.h"

namespace n3ds {

constexpr int TEX_THREADS = 256;

struct TexSample {
    float t[3];        // tint per channel (channel 0 only when C == 1)
    float dx[3], dy[3];  // derivatives w.r.t. the unnormalised pixel coordinates ix, iy
};

// floats per quad entry
template <int C>
constexpr int quad_floats() { return C == 3 ? 16 : 4; }

// Bilinear tap of a packed quad texture at texcoords (u along W, v along H)
// in [0, 1]; corners outside the image contribute zero.
template <int C, bool GRAD>
__device__ __forceinline__ TexSample tex_sample(const float4* __restrict__ quads, int H, int W, float u, float v) {
    // the reference maps [0, 1] to [-1, 1] (uvs = texcoords * 2 - 1) and
    // grid_sample maps that back to pixels ((g + 1) / 2) * (size - 1)
    const float gu = __fsub_rn(__fmul_rn(u, 2.f), 1.f), gv = __fsub_rn(__fmul_rn(v, 2.f), 1.f);
    const float ix = __fmul_rn(__fadd_rn(gu, 1.f) * 0.5f, (float)(W - 1));
    const float iy = __fmul_rn(__fadd_rn(gv, 1.f) * 0.5f, (float)(H - 1));
    // a floor outside [-2, size] leaves every corner outside (NaN lands there too),
    // and keeps the integer corner indices far from overflow
    const float fx = fminf(fmaxf(floorf(ix), -2.f), (float)W), fy = fminf(fmaxf(floorf(iy), -2.f), (float)H);
    const int x0 = (int)fx, y0 = (int)fy;
    TexSample s;
#pragma unroll
    for (int c = 0; c < 3; c++) s.t[c] = s.dx[c] = s.dy[c] = 0.f;
    if (x0 < -1 || x0 > W - 1 || y0 < -1 || y0 > H - 1) return s;  // no corner on the image
    const float ex = __fsub_rn(fx + 1.f, ix), ox = __fsub_rn(ix, fx);  // x1 - ix, ix - x0
    const float ey = __fsub_rn(fy + 1.f, iy), oy = __fsub_rn(iy, fy);
    // corner order nw, ne, sw, se
    const float w[4] = {__fmul_rn(ex, ey), __fmul_rn(ox, ey), __fmul_rn(ex, oy), __fmul_rn(ox, oy)};
    const float wdx[4] = {-ey, ey, -oy, oy}, wdy[4] = {-ex, -ox, ex, ox};
    const float4* q = quads + ((int64_t)(y0 + 1) * (W + 1) + (x0 + 1)) * (quad_floats<C>() / 4);
    float val[4 * C];
    if (C == 3) {
        const float4 a = q[0], b = q[1], c = q[2];
        const float t[12] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w, c.x, c.y, c.z, c.w};
#pragma unroll
        for (int k = 0; k < 4 * C; k++) val[k] = t[k];
    } else {
        const float4 a = q[0];
        const float t[4] = {a.x, a.y, a.z, a.w};
#pragma unroll
        for (int k = 0; k < 4 * C; k++) val[k] = t[k];
    }
#pragma unroll
    for (int k = 0; k < 4; k++) {
#pragma unroll
        for (int c = 0; c < C; c++) {
            s.t[c] = __fadd_rn(s.t[c], __fmul_rn(val[k * C + c], w[k]));
            if (GRAD) {
                s.dx[c] += val[k * C + c] * wdx[k];
                s.dy[c] += val[k * C + c] * wdy[k];
            }
        }
    }
    return s;
}

// quads[(j * (W + 1) + i) * F + k * C + c] = image[c][j - 1 + dy_k][i - 1 + dx_k] or 0 off the image
template <int C>
__global__ __launch_bounds__(TEX_THREADS) void texture_pack_kernel(float* __restrict__ quads,
                                                                   const float* __restrict__ image, int H, int W) {
    constexpr int F = quad_floats<C>();
    const int64_t n = (int64_t)(H + 1) * (W + 1);
    for (int64_t e = (int64_t)blockIdx.x * TEX_THREADS + threadIdx.x; e < n; e += (int64_t)gridDim.x * TEX_THREADS) {
        const int j = (int)(e / (W + 1)), i = (int)(e % (W + 1));
        float out[F];
#pragma unroll
        for (int k = 0; k < F; k++) out[k] = 0.f;
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const int x = i - 1 + (k & 1), y = j - 1 + (k >> 1);
            if (x < 0 || x >= W || y < 0 || y >= H) continue;
#pragma unroll
            for (int c = 0; c < C; c++) out[k * C + c] = image[((int64_t)c * H + y) * W + x];
        }
        float4* dst = reinterpret_cast<float4*>(quads + e * F);
#pragma unroll
        for (int k = 0; k < F / 4; k++) dst[k] = make_float4(out[4 * k], out[4 * k + 1], out[4 * k + 2], out[4 * k + 3]);
    }
}

// the forward streams its per-pair data non-temporally, so it does not displace the
// packed texture from L2 (1.99 -> 1.87 ms on 131M pairs; the same in the backward
// measured slower)
__device__ __forceinline__ float2 ntload2(const float* p, int64_t i) {
    return make_float2(__builtin_nontemporal_load(p + 2 * i), __builtin_nontemporal_load(p + 2 * i + 1));
}

template <int C>
__device__ __forceinline__ float tint_mean(const TexSample& s) {
    return C == 3 ? (s.t[0] + s.t[1] + s.t[2]) / 3.f : s.t[0];
}

template <int C, bool MOD>
__global__ __launch_bounds__(TEX_THREADS) void texture_fwd_kernel(float* __restrict__ colors_out,
                                                                  float* __restrict__ alphas_out,
                                                                  const float* __restrict__ texcoords,
                                                                  const float* __restrict__ colors,
                                                                  const float* __restrict__ alphas,
                                                                  const float4* __restrict__ quads, int H, int W,
                                                                  int64_t n, float tr, float one_minus_tr) {
    const int64_t stride = (int64_t)gridDim.x * TEX_THREADS;
    for (int64_t i = (int64_t)blockIdx.x * TEX_THREADS + threadIdx.x; i < n; i += stride) {
        const float2 uv = ntload2(texcoords, i);
        const TexSample s = tex_sample<C, false>(quads, H, W, uv.x, uv.y);
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const float m = __fadd_rn(__fmul_rn(__builtin_nontemporal_load(colors + i * 3 + c), tr), one_minus_tr);
            __builtin_nontemporal_store(__fmul_rn(m, s.t[C == 3 ? c : 0]), colors_out + i * 3 + c);
        }
        if (MOD) __builtin_nontemporal_store(__fmul_rn(__builtin_nontemporal_load(alphas + i), tint_mean<C>(s)),
                                             alphas_out + i);
    }
}

template <int C, bool MOD, bool GUV>
__global__ __launch_bounds__(TEX_THREADS) void texture_bwd_kernel(
    float* __restrict__ g_colors, float* __restrict__ g_alphas, float* __restrict__ g_texcoords,
    const float* __restrict__ g_colors_out, const float* __restrict__ g_alphas_out,
    const float* __restrict__ texcoords, const float* __restrict__ colors, const float* __restrict__ alphas,
    const float4* __restrict__ quads, int H, int W, int64_t n, float tr, float one_minus_tr) {
    const int64_t stride = (int64_t)gridDim.x * TEX_THREADS;
    for (int64_t i = (int64_t)blockIdx.x * TEX_THREADS + threadIdx.x; i < n; i += stride) {
        const float2 uv = reinterpret_cast<const float2*>(texcoords)[i];
        const TexSample s = tex_sample<C, GUV>(quads, H, W, uv.x, uv.y);
        float gt[3] = {0.f, 0.f, 0.f};  // dL/dtint per channel
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const float g = g_colors_out ? g_colors_out[i * 3 + c] : 0.f;
            g_colors[i * 3 + c] = g * s.t[C == 3 ? c : 0] * tr;
            if (GUV) {
                const float m = __fadd_rn(__fmul_rn(colors[i * 3 + c], tr), one_minus_tr);
                gt[C == 3 ? c : 0] += g * m;
            }
        }
        if (MOD) {
            const float ga = g_alphas_out ? g_alphas_out[i] : 0.f;
            g_alphas[i] = ga * tint_mean<C>(s);
            if (GUV) {
                const float gm = ga * alphas[i] / (float)C;
#pragma unroll
                for (int c = 0; c < C; c++) gt[c] += gm;
            }
        }
        if (GUV) {
            float gx = 0.f, gy = 0.f;
#pragma unroll
            for (int c = 0; c < C; c++) {
                gx += gt[c] * s.dx[c];
                gy += gt[c] * s.dy[c];
            }
            // d ix / d u = W - 1 (through u * 2 - 1 and grid_sample's unnormalisation)
            reinterpret_cast<float2*>(g_texcoords)[i] = make_float2(gx * (float)(W - 1), gy * (float)(H - 1));
        }
    }
}

static int tex_check(const float* texcoords, const float* colors, const float* quads, int64_t height,
                     int64_t width, int64_t channels, int64_t n) {
    N3DS_REQUIRE(channels == 1 || channels == 3, "texture should be RGB or grayscale (got %lld channels)",
                 (long long)channels);
    N3DS_REQUIRE(height >= 1 && width >= 1 && height * width <= (int64_t(1) << 26),
                 "texture size %lld x %lld outside [1, 2^26] texels", (long long)height, (long long)width);
    N3DS_REQUIRE(n >= 0, "negative size");
    N3DS_REQUIRE(n == 0 || (texcoords && colors && quads), "null input");
    N3DS_REQUIRE((reinterpret_cast<uintptr_t>(quads) & 15) == 0 && (reinterpret_cast<uintptr_t>(texcoords) & 7) == 0,
                 "packed texture must be 16-byte and texcoords 8-byte aligned");
    return N3DS_OK;
}

static unsigned tex_blocks(int64_t n) {
    // enough blocks to keep every CU busy; the grid-stride loop covers the rest
    return (unsigned)std::min<int64_t>((n + TEX_THREADS - 1) / TEX_THREADS, 256 * 16);
}

}  // namespace n3ds

using namespace n3ds;

N3DS_EXPORT int64_t n3ds_texture_pack_floats(int64_t height, int64_t width, int64_t channels) {
    if (channels != 1 && channels != 3) {
        set_error("texture should be RGB or grayscale (got %lld channels)", (long long)channels);
        return -1;
    }
    if (height < 1 || width < 1 || height * width > (int64_t(1) << 26)) {
        set_error("texture size %lld x %lld outside [1, 2^26] texels", (long long)height, (long long)width);
        return -1;
    }
    return (height + 1) * (width + 1) * (channels == 3 ? 16 : 4);
}

N3DS_EXPORT int n3ds_texture_pack(float* quads, const float* image, int64_t height, int64_t width, int64_t channels,
                                  void* stream) {
    if (n3ds_texture_pack_floats(height, width, channels) < 0) return N3DS_EINVAL;  // error already set
    N3DS_REQUIRE(quads && image, "null argument");
    N3DS_REQUIRE((reinterpret_cast<uintptr_t>(quads) & 15) == 0, "packed texture must be 16-byte aligned");
    const int H = (int)height, W = (int)width;
    const unsigned blocks = tex_blocks((height + 1) * (width + 1));
    if (channels == 3)
        hipLaunchKernelGGL(texture_pack_kernel<3>, dim3(blocks), dim3(TEX_THREADS), 0, (hipStream_t)stream, quads,
                           image, H, W);
    else
        hipLaunchKernelGGL(texture_pack_kernel<1>, dim3(blocks), dim3(TEX_THREADS), 0, (hipStream_t)stream, quads,
                           image, H, W);
    N3DS_LAUNCH_CHECK();
    return N3DS_OK;
}

N3DS_EXPORT int n3ds_texture_forward(float* colors_out, float* alphas_out, const float* texcoords, const float* colors,
                                     const float* alphas, const float* quads, int64_t height, int64_t width,
                                     int64_t channels, int64_t n, double tint_ratio, int modulate_alpha,
                                     void* stream) {
    if (int e = tex_check(texcoords, colors, quads, height, width, channels, n)) return e;
    N3DS_REQUIRE(n == 0 || colors_out, "null output");
    N3DS_REQUIRE(!modulate_alpha || n == 0 || (alphas && alphas_out), "modulate_alpha needs alphas and alphas_out");
    if (n == 0) return N3DS_OK;
    const float tr = (float)tint_ratio, omt = (float)(1.0 - tint_ratio);
    const float4* q = reinterpret_cast<const float4*>(quads);
    const int H = (int)height, W = (int)width;
    auto launch = [&](auto kernel) {
        hipLaunchKernelGGL(kernel, dim3(tex_blocks(n)), dim3(TEX_THREADS), 0, (hipStream_t)stream, colors_out,
                           alphas_out, texcoords, colors, alphas, q, H, W, n, tr, omt);
    };
    if (channels == 3) modulate_alpha ? launch(texture_fwd_kernel<3, true>) : launch(texture_fwd_kernel<3, false>);
    else modulate_alpha ? launch(texture_fwd_kernel<1, true>) : launch(texture_fwd_kernel<1, false>);
    N3DS_LAUNCH_CHECK();
    return N3DS_OK;
}

N3DS_EXPORT int n3ds_texture_backward(float* grad_colors, float* grad_alphas, float* grad_texcoords,
                                      const float* grad_colors_out, const float* grad_alphas_out,
                                      const float* texcoords, const float* colors, const float* alphas,
                                      const float* quads, int64_t height, int64_t width, int64_t channels, int64_t n,
                                      double tint_ratio, int modulate_alpha, void* stream) {
    if (int e = tex_check(texcoords, colors, quads, height, width, channels, n)) return e;
    N3DS_REQUIRE(n == 0 || grad_colors, "null grad_colors");
    N3DS_REQUIRE(!modulate_alpha || n == 0 || (alphas && grad_alphas), "modulate_alpha needs alphas and grad_alphas");
    N3DS_REQUIRE((reinterpret_cast<uintptr_t>(grad_texcoords) & 7) == 0, "grad_texcoords must be 8-byte aligned");
    if (n == 0) return N3DS_OK;
    const float tr = (float)tint_ratio, omt = (float)(1.0 - tint_ratio);
    const float4* q = reinterpret_cast<const float4*>(quads);
    const int H = (int)height, W = (int)width;
    auto launch = [&](auto kernel) {
        hipLaunchKernelGGL(kernel, dim3(tex_blocks(n)), dim3(TEX_THREADS), 0, (hipStream_t)stream, grad_colors,
                           grad_alphas, grad_texcoords, grad_colors_out, grad_alphas_out, texcoords, colors, alphas,
                           q, H, W, n, tr, omt);
    };
    const bool guv = grad_texcoords != nullptr;
    if (channels == 3) {
        if (modulate_alpha) guv ? launch(texture_bwd_kernel<3, true, true>) : launch(texture_bwd_kernel<3, true, false>);
        else guv ? launch(texture_bwd_kernel<3, false, true>) : launch(texture_bwd_kernel<3, false, false>);
    } else {
        if (modulate_alpha) guv ? launch(texture_bwd_kernel<1, true, true>) : launch(texture_bwd_kernel<1, true, false>);
        else guv ? launch(texture_bwd_kernel<1, false, true>) : launch(texture_bwd_kernel<1, false, false>);
    }
    N3DS_LAUNCH_CHECK();
    return N3DS_OK;
}
