// Ordered "over" compositing of strokes: drop-in replacement for the reference's
// compose_forward / compose_backward (source/strokelib/src/compositing.cu:9-157,
// bound in source/strokelib/strokes.py:339,368).
//
// Mapping: one wavefront per point, lanes over strokes in compositing order
// (stroke S-1 first).  "Over" is associative, so transmittance is an
// exclusive product scan across the wavefront and the adjoint's "composite of
// everything behind stroke k" is an affine suffix scan
//     U_k = d_k a_k + (1 - a_k) U_{k+1}
// evaluated with 64-lane shuffles.  The exact adjoint therefore needs neither
// the reference's suffix-by-subtraction nor its division by (1 - a_k):
//     dD/da_k = T_k (d_k - U_{k+1}),
//     dO/da_k = T_k R_{k+1}   with R_{k+1} = prod_{j behind k} (1 - a_j).
// The reference adjoint (compositing.cu:143-155) is kept as a mode.
#include "n3ds_capi.h"
#include "n3ds_math.h"

namespace n3ds {

constexpr int CWAVES = 4;
constexpr int MAXCH = 128;          // transmittance checkpoints kept per point in the adjoint
constexpr int LDS_GDP_MAX = 8192;   // strokes whose density gradients a block sums in LDS

__device__ __forceinline__ float scan_prod_excl(float f, int lane) {
    float v = f;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        float o = __shfl_up(v, off, 64);
        if (lane >= off) v *= o;
    }
    float e = __shfl_up(v, 1, 64);
    return lane == 0 ? 1.f : e;
}

template <int D>
__global__ __launch_bounds__(64 * CWAVES) void compose_fwd_kernel(float* __restrict__ dens_out,
                                                               float* __restrict__ color_out,
                                                               const float* __restrict__ alphas,
                                                               const float* __restrict__ colors,
                                                               const float* __restrict__ dp, int64_t N, int64_t S,
                                                               int64_t ppb) {
    const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // uniform per wavefront
    const int64_t n0 = (int64_t)blockIdx.x * ppb, n1 = min(N, n0 + ppb);
    for (int64_t n = n0 + wave; n < n1; n += CWAVES) {
        float T = 1.f, Dn = 0.f, C[D];
#pragma unroll
        for (int c = 0; c < D; c++) C[c] = 0.f;
        const float* arow = alphas + n * S;
        const float* crow = colors + n * S * D;
        for (int64_t base = 0; base < S; base += 64) {
            int64_t k = S - 1 - (base + lane);
            bool v = k >= 0;
            float a = v ? arow[k] : 0.f;
            float f = 1.f - a;
            float Tk = T * scan_prod_excl(f, lane);
            float w = a * Tk;
            float dk = v ? dp[k] : 0.f;
            Dn += wave_sum(a != 0.f ? dk * w : 0.f);
#pragma unroll
            for (int c = 0; c < D; c++) C[c] += wave_sum(a != 0.f ? crow[(v ? k : 0) * D + c] * w : 0.f);
            T = __shfl(Tk * f, 63, 64);
        }
        if (lane == 0) {
            float sc = 1.f / (1.f + 1e-6f - T);
            dens_out[n] = Dn;
#pragma unroll
            for (int c = 0; c < D; c++) color_out[n * D + c] = clampf(C[c] * sc, 0.f, 1.f);
        }
    }
}

template <int D>
__global__ __launch_bounds__(64 * CWAVES) void compose_bwd_kernel(
    float* __restrict__ g_alpha, float* __restrict__ g_colors, float* __restrict__ g_dp,
    const float* __restrict__ g_dens, const float* __restrict__ g_color, const float* __restrict__ alphas,
    const float* __restrict__ colors, const float* __restrict__ dp, int64_t N, int64_t S, int64_t ppb, int exact) {
    extern __shared__ float lds_gdp[];  // [S] if S <= LDS_GDP_MAX, else unused (global atomics)
    // transmittance entering chunk g * G (pass 1) and entering each chunk of the group being
    // walked back to front (pass 2); G = 1 up to MAXCH chunks, so pass 2 recomputes nothing
    __shared__ float t_in[CWAVES][MAXCH];
    __shared__ float t_grp[CWAVES][MAXCH];
    const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // uniform per wavefront
    const bool lds = S <= LDS_GDP_MAX;
    if (lds)
        for (int64_t i = threadIdx.x; i < S; i += 64 * CWAVES) lds_gdp[i] = 0.f;
    __syncthreads();
    const int64_t n0 = (int64_t)blockIdx.x * ppb, n1 = min(N, n0 + ppb);
    const int nch = (int)((S + 63) / 64);
    const int G = (nch + MAXCH - 1) / MAXCH;
    for (int64_t n = n0 + wave; n < n1; n += CWAVES) {
        const float* arow = alphas + n * S;
        const float* crow = colors + n * S * D;
        // pass 1, front to back: totals and per-chunk incoming transmittance
        float T = 1.f, Dn = 0.f, C[D];
#pragma unroll
        for (int c = 0; c < D; c++) C[c] = 0.f;
        for (int ch = 0; ch < nch; ch++) {
            int64_t k = S - 1 - ((int64_t)ch * 64 + lane);
            bool v = k >= 0;
            float a = v ? arow[k] : 0.f;
            float f = 1.f - a;
            float Tk = T * scan_prod_excl(f, lane);
            float w = a * Tk;
            Dn += wave_sum(a != 0.f ? (v ? dp[k] : 0.f) * w : 0.f);
#pragma unroll
            for (int c = 0; c < D; c++) C[c] += wave_sum(a != 0.f ? crow[(v ? k : 0) * D + c] * w : 0.f);
            if (lane == 0 && ch % G == 0) t_in[wave][ch / G] = T;
            T = __shfl(Tk * f, 63, 64);
        }
        const float Tn = T;
        const float O = 1.f + 1e-6f - Tn;
        const float sc = 1.f / O;
        const float gD = g_dens ? g_dens[n] : 0.f;
        float gC[D];
#pragma unroll
        for (int c = 0; c < D; c++) {
            float s = C[c] * sc;
            gC[c] = (g_color && 0.f <= s && s <= 1.f) ? g_color[n * D + c] : 0.f;
        }
        // pass 2, back to front: affine suffix scan for the composite behind each stroke
        float Ud = 0.f, Uc[D], Rb = 1.f;
#pragma unroll
        for (int c = 0; c < D; c++) Uc[c] = 0.f;
        for (int ch = nch - 1; ch >= 0; ch--) {
            if (G > 1 && (ch % G == G - 1 || ch == nch - 1)) {
                // entering a group from behind: its chunks' incoming transmittance, recomputed
                // front to back from its checkpoint exactly as pass 1 formed it
                const int g0 = (ch / G) * G;
                float Tg = t_in[wave][ch / G];
                for (int c2 = g0; c2 <= ch; c2++) {
                    const int64_t k2 = S - 1 - ((int64_t)c2 * 64 + lane);
                    const float f2 = 1.f - (k2 >= 0 ? arow[k2] : 0.f);
                    if (lane == 0) t_grp[wave][c2 - g0] = Tg;
                    Tg = __shfl(Tg * scan_prod_excl(f2, lane) * f2, 63, 64);
                }
            }
            int64_t k = S - 1 - ((int64_t)ch * 64 + lane);
            bool v = k >= 0;
            float a = v ? arow[k] : 0.f;
            float f = 1.f - a;
            float dk = v ? dp[k] : 0.f;
            float ck[D];
#pragma unroll
            for (int c = 0; c < D; c++) ck[c] = v ? crow[k * D + c] : 0.f;
            const float Tin = G == 1 ? t_in[wave][ch] : t_grp[wave][ch % G];
            float Tk = Tin * scan_prod_excl(f, lane);
            // inclusive suffix composition (m, bd, bc) of lanes lane..63
            float m = f, bd = dk * a, bc[D];
#pragma unroll
            for (int c = 0; c < D; c++) bc[c] = ck[c] * a;
#pragma unroll
            for (int off = 1; off < 64; off <<= 1) {
                float m2 = __shfl_down(m, off, 64);
                float bd2 = __shfl_down(bd, off, 64);
                float bc2[D];
#pragma unroll
                for (int c = 0; c < D; c++) bc2[c] = __shfl_down(bc[c], off, 64);
                if (lane + off < 64) {
                    bd = bd + m * bd2;
#pragma unroll
                    for (int c = 0; c < D; c++) bc[c] = bc[c] + m * bc2[c];
                    m = m * m2;
                }
            }
            // exclusive (strokes strictly behind this lane) composed with the carry
            float me = __shfl_down(m, 1, 64), bde = __shfl_down(bd, 1, 64), bce[D];
#pragma unroll
            for (int c = 0; c < D; c++) bce[c] = __shfl_down(bc[c], 1, 64);
            if (lane == 63) {
                me = 1.f;
                bde = 0.f;
#pragma unroll
                for (int c = 0; c < D; c++) bce[c] = 0.f;
            }
            float Ubd = bde + me * Ud;       // U_{k+1} (density)
            float Rbk = me * Rb;             // R_{k+1}
            float Ubc[D];
#pragma unroll
            for (int c = 0; c < D; c++) Ubc[c] = bce[c] + me * Uc[c];
            float w = a * Tk;
            float ga;
            if (exact) {
                ga = gD * Tk * (dk - Ubd);
#pragma unroll
                for (int c = 0; c < D; c++) ga += gC[c] * (Tk * (ck[c] - Ubc[c]) * O - C[c] * Tk * Rbk) * (sc * sc);
            } else {
                // compositing.cu:143-154 with T = T_{k+1} and suffix = T_{k+1} U_{k+1}
                float Ta = Tk * f;
                float s1 = 1.f / fmaxf(1.f - a, 1e-6f);
                ga = gD * (Ta * dk - Ta * Ubd) * s1;
                float s2 = s1 * sc * sc;
#pragma unroll
                for (int c = 0; c < D; c++) ga += gC[c] * (Ta * ck[c] * (O - a) - Ta * Ubc[c]) * s2;
            }
            if (v) {
                g_alpha[n * S + k] = ga;
#pragma unroll
                for (int c = 0; c < D; c++) g_colors[(n * S + k) * D + c] = a > 0.f ? gC[c] * sc * w : 0.f;
                if (a > 0.f) atomicAdd(lds ? &lds_gdp[k] : &g_dp[k], gD * w);
            }
            // carry: compose this whole chunk in front of the stroke composite behind it
            float m0 = __shfl(m, 0, 64), bd0 = __shfl(bd, 0, 64);
            Ud = bd0 + m0 * Ud;
#pragma unroll
            for (int c = 0; c < D; c++) Uc[c] = __shfl(bc[c], 0, 64) + m0 * Uc[c];
            Rb = m0 * Rb;
        }
    }
    if (!lds) return;
    __syncthreads();
    for (int64_t i = threadIdx.x; i < S; i += 64 * CWAVES) {
        float v = lds_gdp[i];
        if (v != 0.f) atomicAdd(&g_dp[i], v);
    }
}

static int64_t compose_ppb(int64_t N) {
    int64_t ppb = (N + 4095) / 4096;
    ppb = ppb < CWAVES ? CWAVES : ppb;
    return ((ppb + CWAVES - 1) / CWAVES) * CWAVES;
}

}  // namespace n3ds

using namespace n3ds;

N3DS_EXPORT int n3ds_compose_forward(float* density_out, float* color_out, const float* alphas, const float* colors,
                                     const float* density_params, int64_t n_points, int64_t n_strokes,
                                     int64_t color_dim, void* stream) {
    N3DS_REQUIRE(color_dim == 1 || color_dim == 3, "color_dim must be 1 or 3 (got %lld)", (long long)color_dim);
    N3DS_REQUIRE(n_points >= 0 && n_strokes >= 0, "negative size");
    if (n_points == 0) return N3DS_OK;
    int64_t ppb = compose_ppb(n_points);
    dim3 grid((unsigned)((n_points + ppb - 1) / ppb));
    if (color_dim == 3)
        hipLaunchKernelGGL(compose_fwd_kernel<3>, grid, dim3(64 * CWAVES), 0, (hipStream_t)stream, density_out,
                           color_out, alphas, colors, density_params, n_points, n_strokes, ppb);
    else
        hipLaunchKernelGGL(compose_fwd_kernel<1>, grid, dim3(64 * CWAVES), 0, (hipStream_t)stream, density_out,
                           color_out, alphas, colors, density_params, n_points, n_strokes, ppb);
    N3DS_LAUNCH_CHECK();
    return N3DS_OK;
}

N3DS_EXPORT int n3ds_compose_backward(float* grad_alphas, float* grad_colors, float* grad_density_params,
                                      const float* grad_density_out, const float* grad_color_out,
                                      const float* alphas, const float* colors, const float* density_params,
                                      int64_t n_points, int64_t n_strokes, int64_t color_dim, int adjoint_exact,
                                      void* stream) {
    N3DS_REQUIRE(color_dim == 1 || color_dim == 3, "color_dim must be 1 or 3 (got %lld)", (long long)color_dim);
    N3DS_REQUIRE(n_points >= 0 && n_strokes >= 0, "negative size");
    N3DS_REQUIRE(n_strokes <= (int64_t)64 * MAXCH * MAXCH, "n_strokes (%lld) exceeds %d", (long long)n_strokes,
                 64 * MAXCH * MAXCH);
    if (n_points == 0 || n_strokes == 0) return N3DS_OK;
    int64_t ppb = compose_ppb(n_points);
    dim3 grid((unsigned)((n_points + ppb - 1) / ppb));
    size_t shm = n_strokes <= LDS_GDP_MAX ? (size_t)n_strokes * sizeof(float) : 0;
    if (color_dim == 3)
        hipLaunchKernelGGL(compose_bwd_kernel<3>, grid, dim3(64 * CWAVES), shm, (hipStream_t)stream, grad_alphas,
                           grad_colors, grad_density_params, grad_density_out, grad_color_out, alphas, colors,
                           density_params, n_points, n_strokes, ppb, adjoint_exact);
    else
        hipLaunchKernelGGL(compose_bwd_kernel<1>, grid, dim3(64 * CWAVES), shm, (hipStream_t)stream, grad_alphas,
                           grad_colors, grad_density_params, grad_density_out, grad_color_out, alphas, colors,
                           density_params, n_points, n_strokes, ppb, adjoint_exact);
    N3DS_LAUNCH_CHECK();
    return N3DS_OK;
}
