// Per-stroke transform frame, alpha transfer and the transform adjoint, shared by
// the per-type stroke kernels and the fused stroke-field renderer.
//
// Shape-parameter layout (source/strokelib/strokes.py:245-252): base params,
// then scale (1 or 3), euler xyz (3), translation (3).  The transform is read
// from the end of the row (strokes_sdf.h:164-191).
#pragma once
#include "n3ds_sdf.h"

namespace n3ds {

constexpr uint32_t F_T = 1u, F_R = 2u, F_SS = 4u, F_MS = 8u;

__host__ __device__ constexpr int num_shape_params(uint32_t sdf_id) {
    return num_base_params((int)(sdf_id >> 4)) + ((sdf_id & F_T) ? 3 : 0) + ((sdf_id & F_R) ? 3 : 0) +
           ((sdf_id & F_SS) ? 1 : ((sdf_id & F_MS) ? 3 : 0));
}

struct Frame {
    V3 trans;
    V3 angle;
    Rot R;
    V3 scale;  // SS: replicated
    float sz, cz;  // sin / cos of the z angle (the Euler adjoint's second axis)
    int off_scale, off_angle, off_trans;
};

// row: pointer to the stroke's P shape params
__device__ __forceinline__ Frame load_frame(const float* row, int P, uint32_t flags) {
    Frame F;
    int end = P;
    F.trans = v3s(0.f);
    F.angle = v3s(0.f);
    F.scale = v3s(1.f);
    F.sz = 0.f;
    F.cz = 1.f;
    F.off_trans = F.off_angle = F.off_scale = -1;
    if (flags & F_T) { end -= 3; F.trans = ld3(row + end); F.off_trans = end; }
    if (flags & F_R) {
        end -= 3;
        F.angle = ld3(row + end);
        F.off_angle = end;
        F.R = make_rot(F.angle);
        F.sz = sinf(F.angle.z);
        F.cz = cosf(F.angle.z);
    }
    if (flags & F_SS) { end -= 1; float s = row[end]; F.scale = v3(s, s, s); F.off_scale = end; }
    else if (flags & F_MS) { end -= 3; F.scale = ld3(row + end); F.off_scale = end; }
    return F;
}

// p = Rᵀ (x - t) / s, strokes_sdf.h:164-191 (same operation order)
__device__ __forceinline__ V3 inverse_xform(V3 x, const Frame& F, uint32_t flags) {
    V3 p = x;
    if (flags & F_T) p = p - F.trans;
    if (flags & F_R) p = rot_t(F.R, p);
    if (flags & (F_SS | F_MS)) p = v3(p.x / F.scale.x, p.y / F.scale.y, p.z / F.scale.z);
    return p;
}

__device__ __forceinline__ float alpha_from_sdf(float sdf, float sdf_scale, float delta, bool laplace) {
    if (delta > 0.f) return laplace ? laplace_cdf(-sdf * sdf_scale) : clampf(-sdf * sdf_scale + 0.5f, 0.f, 0.9999f);
    return sdf <= 0.f ? 1.f : 0.f;
}

// dα/dSDF.  Reference mode reads the saved α (strokes_backward.cu:84-93);
// exact mode differentiates the transfer at its argument.
__device__ __forceinline__ float dalpha_dsdf(float sdf, float alpha, float sdf_scale, float delta, bool laplace,
                                             bool exact) {
    if (!(delta > 0.f)) return 0.f;
    if (laplace) {
        if (exact) return -sdf_scale * 0.5f * expf(-fabsf(sdf * sdf_scale));
        return (alpha < 0.5f ? alpha : 1.f - alpha) * -sdf_scale;
    }
    return (0.f < alpha && alpha < 0.9999f) ? -sdf_scale : 0.f;
}
// ∂α/∂(sdf_scale) for the exact radius-scale term
__device__ __forceinline__ float dalpha_dscale(float sdf, float alpha, float sdf_scale, float delta, bool laplace) {
    if (!(delta > 0.f)) return 0.f;
    if (laplace) return -sdf * 0.5f * expf(-fabsf(sdf * sdf_scale));
    return (0.f < alpha && alpha < 0.9999f) ? -sdf : 0.f;
}

// Transform adjoint: given dL/dp (unit space) accumulate scale / euler /
// translation grads and return dL/dx.  q = x - t is recomputed exactly.
__device__ __forceinline__ V3 xform_adjoint(V3 gp, V3 p, V3 x, const Frame& F, uint32_t flags, float& gs1, V3& gs3,
                                            V3& gang, V3& gtr) {
    V3 g = gp;
    if (flags & F_SS) {
        float s = F.scale.x;
        g = v3(g.x / s, g.y / s, g.z / s);
        gs1 += -(g.x * p.x) + -(g.y * p.y) + -(g.z * p.z);
    } else if (flags & F_MS) {
        g = v3(g.x / F.scale.x, g.y / F.scale.y, g.z / F.scale.z);
        gs3 = gs3 + v3(-g.x * p.x, -g.y * p.y, -g.z * p.z);
    }
    if (flags & F_R) {
        V3 q = (flags & F_T) ? x - F.trans : x;
        gang = gang + grad_euler(F.R, F.sz, F.cz, g, q);
        g = rot_n(F.R, g);
    }
    if (flags & F_T) gtr = gtr - g;
    return g;
}

// permute_multiscale_axis (strokes_sdf.h:212-222)
__device__ __forceinline__ V3 permute_ms(V3 d, V3 s) {
    return s.x > s.y ? (s.y > s.z ? v3(d.x, d.y, d.z) : (s.x > s.z ? v3(d.x, d.z, d.y) : v3(d.z, d.x, d.y)))
                     : (s.z > s.y ? v3(d.z, d.y, d.x) : (s.x > s.z ? v3(d.y, d.x, d.z) : v3(d.y, d.z, d.x)));
}

// BaseSDF<...>::texcoord (strokes_sdf.h:258-268, 301-311, 357-367, 456-469, 511-526);
// row points at the stroke's parameter row as the reference's `params` does.
__device__ __forceinline__ void texcoord_eval(int base, V3 p, const float* row, bool ms, float& u, float& v) {
    const float PI_F = 3.14159265358979323846f;
    u = 0.f;
    v = 0.f;
    if (base == SPHERE) {
        V3 d = p * rsqrtf(dot(p, p));
        if (ms) d = permute_ms(d, ld3(row + 0));
        u = asinf(clampf(d.x, -1.f, 1.f)) / PI_F + 0.5f;
        v = atan2f(d.z, d.y) / (2.f * PI_F) + 0.5f;
    } else if (base == CUBE || base == ROUND_CUBE) {
        V3 uvt = p * 0.5f + v3s(0.5f);
        if (ms) uvt = permute_ms(uvt, ld3(row + 0));  // both bases read params + 0
        u = clampf(uvt.x, 0.f, 1.f);
        v = clampf(uvt.y, 0.f, 1.f);
    } else if (base == CAPSULE || base == LINE) {
        float h = row[0];
        if (base == CAPSULE) u = clampf((p.y / (h + 1.f)) * 0.5f + 0.5f, 0.f, 1.f);
        else u = clampf((p.y + h + 1.f - row[1]) / (2.f * h + 2.f), 0.f, 1.f);
        if (ms) {
            V3 s = ld3(row + 1);  // the reference reads params + 1 for both bases
            if (s.z > s.x) p = v3(p.z, p.y, p.x);
        }
        v = atan2f(p.z, p.x) / (2.f * PI_F) + 0.5f;
    }
}

}  // namespace n3ds
