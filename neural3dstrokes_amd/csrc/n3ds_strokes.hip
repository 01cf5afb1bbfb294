// Per-type stroke evaluation: the drop-in replacement for the reference's
// stroke_forward / stroke_backward (source/strokelib/src/strokes_forward.cu,
// strokes_backward.cu; bound in source/strokelib/strokes.py:167,217).
//
// Mapping: one lane per stroke, a wavefront covers 64 consecutive strokes and
// walks a block-private range of points.  The stroke's parameters and its
// rotation are loaded once per lane and kept in registers; the point (x,
// radius, view direction) is wave-uniform and broadcast.  Outputs
// alpha[N,S] / color[N,S,3] are written lane-contiguous.  The adjoint keeps the
// stroke's parameter gradients in registers across the whole point range, adds
// the block's wavefronts together in LDS and issues one atomic per parameter per
// stroke per block (blocks are few and long: about 1024 over all stroke tiles);
// dL/dx is summed across the wavefront before a single atomic per point.
#include "n3ds_capi.h"
#include "n3ds_stroke_common.h"

namespace n3ds {

constexpr int WAVES = 4;
constexpr int BLOCK = 64 * WAVES;

template <int BASE, int CID>
__global__ __launch_bounds__(BLOCK) void stroke_fwd_kernel(
    float* __restrict__ alpha_out, float* __restrict__ color_out, float* __restrict__ sdf_out,
    float* __restrict__ tex_out, const float* __restrict__ x, const float* __restrict__ radius,
    const float* __restrict__ viewdir, const float* __restrict__ sp, const float* __restrict__ cp, int64_t N,
    int64_t S, int64_t spr, int P, int Q, uint32_t flags, float delta, int laplace, int isr, int64_t ppb) {
    constexpr int NB = num_base_params(BASE);
    const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // uniform per wavefront
    const int64_t k = (int64_t)blockIdx.y * 64 + lane;
    const bool kv = k < S;
    const int64_t kk = kv ? k : S - 1;
    const float* row = sp + kk * P;
    float prm[NB > 0 ? NB : 1];
#pragma unroll
    for (int i = 0; i < NB; i++) prm[i] = row[i];
    const float* crow = cp + kk * Q;
    float col[num_color_params(CID)];
#pragma unroll
    for (int i = 0; i < num_color_params(CID); i++) col[i] = crow[i];
    const Frame F = load_frame(row, P, flags);
    const int64_t n0 = (int64_t)blockIdx.x * ppb;
    const int64_t n1 = min(N, n0 + ppb);
    const float kscale = laplace ? 2.f : 0.5f;
    for (int64_t n = n0 + wave; n < n1; n += WAVES) {
        V3 xw = ld3(x + 3 * n);
        float r = radius[n];
        V3 dir = (CID >= 2) ? ld3(viewdir + 3 * (n / spr)) : v3s(0.f);
        V3 p = inverse_xform(xw, F, flags);
        float sdf = sdf_eval(BASE, p, prm);
        if (isr) {
            if (flags & F_SS) r = r / F.scale.x;
            else if (flags & F_MS) r = r / radius_factor(ref_dir(BASE, p, prm), F.scale);
        }
        float sc = kscale / (delta * r);
        float a = alpha_from_sdf(sdf, sc, delta, laplace);
        V3 c = color_eval_t<CID>(p, dir, col);
        if (kv) {
            int64_t o = n * S + k;
            alpha_out[o] = a;
            color_out[3 * o + 0] = c.x;
            color_out[3 * o + 1] = c.y;
            color_out[3 * o + 2] = c.z;
            if (sdf_out) sdf_out[o] = sdf;
            if (tex_out) {
                float u, v;
                texcoord_eval(BASE, p, row, (flags & F_MS) != 0, u, v);
                tex_out[2 * o + 0] = u;
                tex_out[2 * o + 1] = v;
            }
        }
    }
}

template <int BASE, int CID>
__global__ __launch_bounds__(BLOCK) void stroke_bwd_kernel(
    float* __restrict__ gshape, float* __restrict__ gcolor, float* __restrict__ gx,
    const float* __restrict__ galpha, const float* __restrict__ gcol_in, const float* __restrict__ gsdf,
    const float* __restrict__ x, const float* __restrict__ radius, const float* __restrict__ viewdir,
    const float* __restrict__ alpha, const float* __restrict__ sp, const float* __restrict__ cp, int64_t N,
    int64_t S, int64_t spr, int P, int Q, uint32_t flags, float delta, int laplace, int isr, int exact,
    int64_t ppb) {
    constexpr int NB = num_base_params(BASE);
    constexpr int NQ = num_color_params(CID);
    const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // uniform per wavefront
    const int64_t k = (int64_t)blockIdx.y * 64 + lane;
    const bool kv = k < S;
    const int64_t kk = kv ? k : S - 1;
    const float* row = sp + kk * P;
    float prm[NB > 0 ? NB : 1];
#pragma unroll
    for (int i = 0; i < NB; i++) prm[i] = row[i];
    const float* crow = cp + kk * Q;
    float col[NQ];
#pragma unroll
    for (int i = 0; i < NQ; i++) col[i] = crow[i];
    const Frame F = load_frame(row, P, flags);
    float gb[NB > 0 ? NB : 1];
#pragma unroll
    for (int i = 0; i < (NB > 0 ? NB : 1); i++) gb[i] = 0.f;
    float gc[NQ];
#pragma unroll
    for (int i = 0; i < NQ; i++) gc[i] = 0.f;
    float gs1 = 0.f;
    V3 gs3 = v3s(0.f), gang = v3s(0.f), gtr = v3s(0.f);
    const int64_t n0 = (int64_t)blockIdx.x * ppb;
    const int64_t n1 = min(N, n0 + ppb);
    const float kscale = laplace ? 2.f : 0.5f;
    for (int64_t n = n0 + wave; n < n1; n += WAVES) {
        const int64_t o = n * S + kk;
        V3 xw = ld3(x + 3 * n);
        float r = radius[n];
        V3 dir = (CID >= 2) ? ld3(viewdir + 3 * (n / spr)) : v3s(0.f);
        V3 p = inverse_xform(xw, F, flags);
        float ga = kv ? galpha[o] : 0.f;
        V3 gcv = kv ? ld3(gcol_in + 3 * o) : v3s(0.f);
        float gsd = (kv && gsdf) ? gsdf[o] : 0.f;
        float fac = 1.f;
        if (isr) {
            if (flags & F_SS) r = r / F.scale.x;
            else if (flags & F_MS) { fac = radius_factor(ref_dir(BASE, p, prm), F.scale); r = r / fac; }
        }
        float sc = kscale / (delta * r);
        float sdf = 0.f;
        float a;
        if (exact) {
            sdf = sdf_eval(BASE, p, prm);
            a = alpha_from_sdf(sdf, sc, delta, laplace);
        } else {
            a = kv ? alpha[o] : 0.f;
        }
        float gsdf_tot = ga * dalpha_dsdf(sdf, a, sc, delta, laplace, exact) + gsd;
        V3 gp = color_grad_t<CID>(p, dir, col, gcv, gc, exact != 0);
        V3 dsdp;
        if (exact) dsdp = grad_ex(BASE, p, prm, gsdf_tot, gb);
        else dsdp = grad_ref(BASE, p, prm, gsdf_tot, gb);
        gp = gp + dsdp * gsdf_tot;
        if (exact && isr && (flags & (F_SS | F_MS))) {
            float gsc = ga * dalpha_dscale(sdf, a, sc, delta, laplace);
            if (flags & F_SS) {
                gs1 += gsc * sc / F.scale.x;
            } else {
                radius_factor_adjoint(BASE, p, prm, F.scale, gsc * sc / fac, gp, gs3, gb);
            }
        }
        V3 gxv = xform_adjoint(gp, p, xw, F, flags, gs1, gs3, gang, gtr);
        if (gx) {
            if (!kv) gxv = v3s(0.f);
            float sx = wave_sum(gxv.x), sy = wave_sum(gxv.y), sz = wave_sum(gxv.z);
            if (lane == 0) {
                atomicAdd(gx + 3 * n + 0, sx);
                atomicAdd(gx + 3 * n + 1, sy);
                atomicAdd(gx + 3 * n + 2, sz);
            }
        }
    }
    // the block's wavefronts hold partial sums for the same 64 strokes: add them in LDS,
    // then the first wavefront issues one global atomic per parameter per stroke
    constexpr int NR = (NB > 0 ? NB : 1) + 10 + NQ;
    __shared__ float red[NR * 64];
    for (int i = threadIdx.x; i < NR * 64; i += BLOCK) red[i] = 0.f;
    __syncthreads();
    {
        float v[NR];
#pragma unroll
        for (int i = 0; i < (NB > 0 ? NB : 1); i++) v[i] = NB > 0 ? gb[i] : 0.f;
        constexpr int o = NB > 0 ? NB : 1;
        v[o] = gs1;
        v[o + 1] = gs3.x; v[o + 2] = gs3.y; v[o + 3] = gs3.z;
        v[o + 4] = gang.x; v[o + 5] = gang.y; v[o + 6] = gang.z;
        v[o + 7] = gtr.x; v[o + 8] = gtr.y; v[o + 9] = gtr.z;
#pragma unroll
        for (int i = 0; i < NQ; i++) v[o + 10 + i] = gc[i];
#pragma unroll
        for (int i = 0; i < NR; i++) atomicAdd(&red[i * 64 + lane], v[i]);
    }
    __syncthreads();
    if (wave != 0 || !kv) return;
    const float* r = red + lane;
    constexpr int o = NB > 0 ? NB : 1;
    float* grow = gshape + k * P;
#pragma unroll
    for (int i = 0; i < NB; i++) atomicAdd(grow + i, r[i * 64]);
    if (F.off_scale >= 0) {
        if (flags & F_SS) atomicAdd(grow + F.off_scale, r[o * 64]);
        else for (int i = 0; i < 3; i++) atomicAdd(grow + F.off_scale + i, r[(o + 1 + i) * 64]);
    }
    if (F.off_angle >= 0)
        for (int i = 0; i < 3; i++) atomicAdd(grow + F.off_angle + i, r[(o + 4 + i) * 64]);
    if (F.off_trans >= 0)
        for (int i = 0; i < 3; i++) atomicAdd(grow + F.off_trans + i, r[(o + 7 + i) * 64]);
    float* gcrow = gcolor + k * Q;
#pragma unroll
    for (int i = 0; i < NQ; i++) atomicAdd(gcrow + i, r[(o + 10 + i) * 64]);
}

typedef void (*FwdFn)(float*, float*, float*, float*, const float*, const float*, const float*, const float*,
                      const float*, int64_t, int64_t, int64_t, int, int, uint32_t, float, int, int, int64_t);
typedef void (*BwdFn)(float*, float*, float*, const float*, const float*, const float*, const float*,
                      const float*, const float*, const float*, const float*, const float*, int64_t, int64_t,
                      int64_t, int, int, uint32_t, float, int, int, int, int64_t);

template <int B, int C>
struct Tab {
    static constexpr FwdFn f = &stroke_fwd_kernel<B, C>;
    static constexpr BwdFn b = &stroke_bwd_kernel<B, C>;
};

#define N3DS_ROW(B) {Tab<B, 0>::f, Tab<B, 1>::f, Tab<B, 2>::f, Tab<B, 3>::f}
#define N3DS_ROWB(B) {Tab<B, 0>::b, Tab<B, 1>::b, Tab<B, 2>::b, Tab<B, 3>::b}
static const FwdFn kFwd[NUM_BASE][4] = {N3DS_ROW(0), N3DS_ROW(1), N3DS_ROW(2), N3DS_ROW(3), N3DS_ROW(4), N3DS_ROW(5),
                                        N3DS_ROW(6), N3DS_ROW(7), N3DS_ROW(8), N3DS_ROW(9), N3DS_ROW(10), N3DS_ROW(11)};
static const BwdFn kBwd[NUM_BASE][4] = {N3DS_ROWB(0), N3DS_ROWB(1), N3DS_ROWB(2), N3DS_ROWB(3), N3DS_ROWB(4), N3DS_ROWB(5),
                                        N3DS_ROWB(6), N3DS_ROWB(7), N3DS_ROWB(8), N3DS_ROWB(9), N3DS_ROWB(10), N3DS_ROWB(11)};

// points per block for about target_blocks blocks over the stroke tiles (the adjoint
// uses fewer, longer blocks: each block ends with an atomic per stroke parameter)
static int64_t points_per_block(int64_t N, int64_t S, int64_t target_blocks = 4096) {
    int64_t tiles = (S + 63) / 64;
    int64_t ppb = (N * tiles + target_blocks - 1) / target_blocks;
    ppb = ppb < 64 ? 64 : (ppb > 16384 ? 16384 : ppb);
    return ((ppb + WAVES - 1) / WAVES) * WAVES;
}

static int check_common(int64_t N, int64_t R, int64_t S, int64_t P, int64_t Q, uint32_t sdf_id, uint32_t color_id) {
    N3DS_REQUIRE(N >= 0 && S >= 0 && R >= 0, "negative size");
    N3DS_REQUIRE((sdf_id >> 4) < (uint32_t)NUM_BASE, "sdf_id %u out of range", sdf_id);
    N3DS_REQUIRE(color_id < 4u, "color_id %u out of range", color_id);
    N3DS_REQUIRE(P == num_shape_params(sdf_id), "shape_params has %lld columns, sdf_id %u needs %d", (long long)P,
                 sdf_id, num_shape_params(sdf_id));
    N3DS_REQUIRE(Q == num_color_params((int)color_id), "color_params has %lld columns, color_id %u needs %d",
                 (long long)Q, color_id, num_color_params((int)color_id));
    N3DS_REQUIRE(N == 0 || (R > 0 && N % R == 0), "n_points (%lld) must be a multiple of n_viewdirs (%lld)",
                 (long long)N, (long long)R);
    N3DS_REQUIRE((uint32_t)((sdf_id & F_SS) != 0) + (uint32_t)((sdf_id & F_MS) != 0) <= 1u,
                 "sdf_id %u sets both single and multi scale", sdf_id);
    return N3DS_OK;
}

}  // namespace n3ds

using namespace n3ds;

N3DS_EXPORT int n3ds_stroke_forward(float* alpha_out, float* color_out, float* sdf_out, float* texcoord_out,
                                    const float* x, const float* radius, const float* viewdir,
                                    const float* shape_params, const float* color_params, int64_t n_points,
                                    int64_t n_viewdirs, int64_t n_strokes, int64_t n_shape_params,
                                    int64_t n_color_params, uint32_t sdf_id, uint32_t color_id, float sdf_delta,
                                    int use_laplace_transform, int inv_scale_radius, void* stream) {
    int st = check_common(n_points, n_viewdirs, n_strokes, n_shape_params, n_color_params, sdf_id, color_id);
    if (st) return st;
    if (n_points == 0 || n_strokes == 0) return N3DS_OK;
    int64_t ppb = points_per_block(n_points, n_strokes);
    dim3 grid((unsigned)((n_points + ppb - 1) / ppb), (unsigned)((n_strokes + 63) / 64));
    hipLaunchKernelGGL(kFwd[sdf_id >> 4][color_id], grid, dim3(BLOCK), 0, (hipStream_t)stream, alpha_out, color_out,
                       sdf_out, texcoord_out, x, radius, viewdir, shape_params, color_params, n_points, n_strokes,
                       n_points / n_viewdirs, (int)n_shape_params, (int)n_color_params, sdf_id & 15u, sdf_delta,
                       use_laplace_transform, inv_scale_radius, ppb);
    N3DS_LAUNCH_CHECK();
    return N3DS_OK;
}

N3DS_EXPORT int n3ds_stroke_backward(float* grad_shape_params, float* grad_color_params, float* grad_x,
                                     const float* grad_alpha, const float* grad_color, const float* grad_sdf,
                                     const float* x, const float* radius, const float* viewdir, const float* alpha,
                                     const float* shape_params, const float* color_params, int64_t n_points,
                                     int64_t n_viewdirs, int64_t n_strokes, int64_t n_shape_params,
                                     int64_t n_color_params, uint32_t sdf_id, uint32_t color_id, float sdf_delta,
                                     int use_laplace_transform, int inv_scale_radius, int adjoint_exact,
                                     void* stream) {
    int st = check_common(n_points, n_viewdirs, n_strokes, n_shape_params, n_color_params, sdf_id, color_id);
    if (st) return st;
    if (n_points == 0 || n_strokes == 0) return N3DS_OK;
    int64_t ppb = points_per_block(n_points, n_strokes, 1024);
    dim3 grid((unsigned)((n_points + ppb - 1) / ppb), (unsigned)((n_strokes + 63) / 64));
    hipLaunchKernelGGL(kBwd[sdf_id >> 4][color_id], grid, dim3(BLOCK), 0, (hipStream_t)stream, grad_shape_params,
                       grad_color_params, grad_x, grad_alpha, grad_color, grad_sdf, x, radius, viewdir, alpha,
                       shape_params, color_params, n_points, n_strokes, n_points / n_viewdirs, (int)n_shape_params,
                       (int)n_color_params, sdf_id & 15u, sdf_delta, use_laplace_transform, inv_scale_radius,
                       adjoint_exact, ppb);
    N3DS_LAUNCH_CHECK();
    return N3DS_OK;
}
