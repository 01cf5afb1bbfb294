// Stroke signed-distance functions, colour fields and their adjoints (gfx950).
//
// Unit-space base SDFs follow source/strokelib/src/strokes_sdf.h:242-1456;
// the colour fields follow strokes_color.h:34-171 and sh.h:8-30; transforms
// follow strokes_sdf.h:25-222.  Every base type has three entry points:
//
//   eval      forward value
//   grad_ex   exact adjoint: returns dSDF/dp and adds g * dSDF/dparam into a
//             local array (the true derivative of eval, conventions sign(0)=+1,
//             d|v|/dv = 0 at v = 0)
//   grad_ref  the reference's hand-written adjoint (strokes_sdf.h grad_sdf
//             members), kept term for term including where it differs from the
//             derivative (cube outside, capped torus, quadratic Bezier comma
//             operators, curve squared-distance gradients).  Like the reference
//             it returns the value the caller multiplies by dL/dSDF.
#pragma once
#include "n3ds_math.h"

namespace n3ds {

enum BaseType : int {
    SPHERE = 0, CUBE = 1, ROUND_CUBE = 2, CAPPED_TORUS = 3, CAPSULE = 4, LINE = 5, TRIPRISM = 6,
    OCTAHEDRON = 7, TETRAHEDRON = 8, QUAD_BEZIER = 9, CUBIC_BEZIER = 10, CATMULL_ROM = 11, NUM_BASE = 12
};

__host__ __device__ constexpr int num_base_params(int b) {
    return b == ROUND_CUBE ? 1 : b == CAPPED_TORUS ? 2 : b == CAPSULE ? 1 : b == LINE ? 2 : b == TRIPRISM ? 1
         : b == QUAD_BEZIER ? 11 : (b == CUBIC_BEZIER || b == CATMULL_ROM) ? 14 : 0;
}
__host__ __device__ constexpr bool is_curve(int b) { return b >= QUAD_BEZIER; }

constexpr float OCTA_C = 0.57735027f;
constexpr float TRI_C = 0.866025f;
constexpr int CURVE_SEGMENTS = 10;

// ---------------------------------------------------------------------------
// rotation R = Rz(z) Ry(y) Rx(x), strokes_sdf.h:28-39
// ---------------------------------------------------------------------------
struct Rot {
    float r[3][3];
};

__device__ __forceinline__ Rot make_rot(V3 a) {
    float sx = sinf(a.x), sy = sinf(a.y), sz = sinf(a.z);
    float cx = cosf(a.x), cy = cosf(a.y), cz = cosf(a.z);
    Rot R;
    R.r[0][0] = cy * cz;
    R.r[0][1] = sx * sy * cz - cx * sz;
    R.r[0][2] = sx * sz + cx * sy * cz;
    R.r[1][0] = cy * sz;
    R.r[1][1] = cx * cz + sx * sy * sz;
    R.r[1][2] = cx * sy * sz - sx * cz;
    R.r[2][0] = -sy;
    R.r[2][1] = sx * cy;
    R.r[2][2] = cx * cy;
    return R;
}
// Rᵀ p  (rotate_point<true>)
__device__ __forceinline__ V3 rot_t(const Rot& R, V3 p) {
    return v3(R.r[0][0] * p.x + R.r[1][0] * p.y + R.r[2][0] * p.z,
              R.r[0][1] * p.x + R.r[1][1] * p.y + R.r[2][1] * p.z,
              R.r[0][2] * p.x + R.r[1][2] * p.y + R.r[2][2] * p.z);
}
// R p  (rotate_point<false>, and grad_point_rotate_point<true>)
__device__ __forceinline__ V3 rot_n(const Rot& R, V3 p) {
    return v3(R.r[0][0] * p.x + R.r[0][1] * p.y + R.r[0][2] * p.z,
              R.r[1][0] * p.x + R.r[1][1] * p.y + R.r[1][2] * p.z,
              R.r[2][0] * p.x + R.r[2][1] * p.y + R.r[2][2] * p.z);
}
// dL/d(euler) for p = Rᵀ q given g = dL/dp.  With w = (R g) × q:
// dL/dx = R[:,0]·w, dL/dy = (-sin z, cos z, 0)·w, dL/dz = w.z.  (Closed form of
// sum_ij g_j q_i dR_ij/da for R = Rz Ry Rx.)
// sz, cz: sin / cos of the z angle, computed once per stroke.
__device__ __forceinline__ V3 grad_euler(const Rot& R, float sz, float cz, V3 g, V3 q) {
    V3 Rg = rot_n(R, g);
    V3 w = cross(Rg, q);
    return v3(R.r[0][0] * w.x + R.r[1][0] * w.y + R.r[2][0] * w.z, -sz * w.x + cz * w.y, w.z);
}

// ---------------------------------------------------------------------------
// primitives: forward values
// ---------------------------------------------------------------------------
__device__ __forceinline__ float eval_sphere(V3 p) {
    return sqrtf(p.x * p.x + p.y * p.y + p.z * p.z + 1e-8f) - 1.f;
}
__device__ __forceinline__ float eval_cube(V3 p) {
    V3 q = vabs(p) - v3s(1.f);
    return safe_len(vmax0(q)) + fminf(fmaxf(q.x, fmaxf(q.y, q.z)), 0.f);
}
__device__ __forceinline__ float eval_round_cube(V3 p, const float* prm) {
    V3 d = vabs(p) - v3s(1.f);
    V3 dp = vmax0(d);
    float n = sqrtf(dp.x * dp.x + dp.y * dp.y + dp.z * dp.z + 1e-8f);
    return n + fminf(fmaxf(d.x, fmaxf(d.y, d.z)), 0.f) - prm[0];
}
__device__ __forceinline__ float eval_capped_torus(V3 p, const float* prm) {
    float ax = fabsf(p.x);
    float s = sinf(prm[0]), c = cosf(prm[0]);
    float k = (ax * c - p.y * s) > 0.f ? ax * s + p.y * c : sqrtf(ax * ax + p.y * p.y + 1e-8f);
    return sqrtf(ax * ax + p.y * p.y + p.z * p.z + 1.f - 2.f * k + 1e-8f) - prm[1];
}
__device__ __forceinline__ float eval_capsule(V3 p, const float* prm) {
    float h = prm[0];
    float y = p.y - clampf(p.y, -h, h);
    return sqrtf(p.x * p.x + y * y + p.z * p.z + 1e-8f) - 1.f;
}
__device__ __forceinline__ float eval_line(V3 p, const float* prm) {
    float h = prm[0], rd = prm[1];
    float t = clampf((p.y + h) / (2.f * h), 0.f, 1.f);
    float r = lerpf(1.f - rd, 1.f + rd, t);
    float y = p.y - clampf(p.y, -h, h);
    return sqrtf(p.x * p.x + y * y + p.z * p.z + 1e-8f) - r;
}
__device__ __forceinline__ float eval_triprism(V3 p, const float* prm) {
    float qx = fabsf(p.x), qy = fabsf(p.y);
    return fmaxf(qy - prm[0], fmaxf(qx * TRI_C + p.z * 0.5f, -p.z) - 0.5f);
}
__device__ __forceinline__ float eval_octahedron(V3 p) {
    return (fabsf(p.x) + fabsf(p.y) + fabsf(p.z) - 1.f) * OCTA_C;
}
__device__ __forceinline__ float eval_tetrahedron(V3 p) {
    return (fmaxf(fabsf(p.x + p.y) - p.z, fabsf(p.x - p.y) + p.z) - 1.f) * OCTA_C;
}

// ---------------------------------------------------------------------------
// quadratic Bezier tube (closed-form cubic solve), strokes_sdf.h:647-705
//
// The closed form picks t from the depressed cubic s^3 + 3 p s + q = 0
// (s = t + kx), whose coefficients are differences of nearly equal terms.
// The set-up (kx, ky, kz, p, q, h) is therefore computed in double (a few
// dozen FP64 operations, the unpacked FP32 rate on CDNA4), the transcendentals
// in single precision.  With its 1e-8 regularisers the closed form does not
// return a root of that cubic but the exact root of a nearby one:
//   one-root branch:  s^3 + 3 p' s + q = 0,  p' = cbrt(p^3 + 2.5e-9)
//     (the two cube roots u, w satisfy u^3 + w^3 = -q, u w = -p');
//   three-root branch, second candidate:  s^3 + 3 (p - 1e-8) s + q'' = 0,
//     q'' = q (p - 1e-8) / (p + 5e-9 / z),  z = sqrt(1e-8 - p)
//     (s = 2 z cos(v + 2 pi / 3) with cos 3v = q / (2 p z + 1e-8)).
// For those candidates t is refined by one Newton step on that cubic, written
// as F(t) = g(t)/|b|^2 + c1 (t + kx) + c0 with the geometric residual
// g(t) = (P(t) - pos) . P'(t)/2 (no cancellation) and the small regulariser
// terms c1, c0 in closed form; the exact adjoint differentiates F = 0
// implicitly.  The three-root branch's first candidate (m + n) z - kx is not a
// root of any of these cubics and keeps the closed form and its derivative.
// ---------------------------------------------------------------------------
struct QuadSolve {
    V3 a, b, c, d;
    float kk, kx, ky, kz, p, q;
    float z, v, m, n;   // three-root branch
    float den, sq;      // acos argument's denominator, sqrt(1 - arg^2)
    float c1, c0;       // regulariser terms of F
    float ep, eq;       // dC/dp (one-root: per unit of s), dC/dq of C = c1 (t + kx) + c0
    bool one;           // h >= 0
    bool use_b;         // three-root branch picked the second candidate
    bool clamped;       // acos argument outside (-1, 1)
    bool root;          // t is an interior root of F (refined; implicit adjoint)
    float traw;         // unclamped parameter of the pick (outside (0, 1) when clamped)
    float t;            // chosen parameter (clamped)
    float dist2;        // |P(t) - pos|^2
};

__device__ __forceinline__ float cbrt_signed(float x) { return copysignf(powf(fabsf(x), 1.0f / 3.0f), x); }

__device__ __forceinline__ float quad_d2(V3 d, V3 c, V3 b, float t) {
    V3 m = d + (c + b * t) * t;
    return dot(m, m);
}

__device__ __forceinline__ QuadSolve quad_solve(V3 pos, const float* prm) {
    QuadSolve Q;
    const double Ax = prm[0], Ay = prm[1], Az = prm[2];
    const double Bx = prm[3], By = prm[4], Bz = prm[5];
    const double Cx = prm[6], Cy = prm[7], Cz = prm[8];
    const double ax = Bx - Ax, ay = By - Ay, az = Bz - Az;
    const double bx = Ax - 2.0 * Bx + Cx, by = Ay - 2.0 * By + Cy, bz = Az - 2.0 * Bz + Cz;
    const double dx = Ax - (double)pos.x, dy = Ay - (double)pos.y, dz = Az - (double)pos.z;
    const double kk = 1.0 / (bx * bx + by * by + bz * bz);
    const double kx = kk * (ax * bx + ay * by + az * bz);
    const double ky = kk * (2.0 * (ax * ax + ay * ay + az * az) + (dx * bx + dy * by + dz * bz)) / 3.0;
    const double kz = kk * (dx * ax + dy * ay + dz * az);
    const double p = ky - kx * kx;
    const double q = kx * (2.0 * kx * kx - 3.0 * ky) + kz;
    const double h = q * q + 4.0 * p * p * p;
    Q.a = v3((float)ax, (float)ay, (float)az);
    Q.b = v3((float)bx, (float)by, (float)bz);
    Q.c = Q.a * 2.f;
    Q.d = v3((float)dx, (float)dy, (float)dz);
    Q.kk = (float)kk;
    Q.kx = (float)kx;
    Q.ky = (float)ky;
    Q.kz = (float)kz;
    Q.p = (float)p;
    Q.q = (float)q;
    Q.one = h >= 0.0;
    Q.use_b = false;
    Q.clamped = false;
    Q.z = Q.v = Q.m = Q.n = Q.den = Q.sq = 0.f;
    float traw;
    if (Q.one) {
        const double hh = sqrt(h + 1e-8);
        traw = (cbrt_signed((float)((hh - q) * 0.5)) + cbrt_signed((float)(-(hh + q) * 0.5))) - Q.kx;
        // p' - p = 2.5e-9 / (p'^2 + p' p + p^2), free of cancellation
        const float pp = cbrt_signed((float)(p * p * p + 2.5e-9));
        const float dl = 2.5e-9f / (pp * pp + pp * Q.p + Q.p * Q.p);
        Q.c1 = 3.f * dl;
        Q.c0 = 0.f;
        Q.ep = pp != 0.f ? -3.f * dl * (pp + Q.p) / (pp * pp) : 0.f;  // d(3 dl)/dp
        Q.eq = 0.f;
    } else {
        const double z = sqrt(-p + 1e-8);
        const double den = p * z * 2.0 + 1e-8;
        const double arg = q / den;
        // the regularisers can push |arg| past 1 when h ~ 0, where the reference's
        // acosf returns NaN; acos is taken of the argument clamped to [-1, 1]
        Q.clamped = !(arg > -1.0 && arg < 1.0);
        // acos via asin of the half complement, which double precision keeps exact
        const float om = (float)fmax(1.0 - arg, 0.0), op = (float)fmax(1.0 + arg, 0.0);
        const float ac = arg >= 0.0 ? 2.f * asinf(fminf(sqrtf(om * 0.5f), 1.f))
                                    : 3.14159265358979f - 2.f * asinf(fminf(sqrtf(op * 0.5f), 1.f));
        Q.z = (float)z;
        Q.den = (float)den;
        Q.sq = sqrtf(om * op);
        Q.v = ac / 3.f;
        Q.m = cosf(Q.v);
        Q.n = sinf(Q.v) * 1.732050808f;
        const float ra = (Q.m + Q.n) * Q.z - Q.kx, rb = (-Q.n - Q.m) * Q.z - Q.kx;
        const float da = quad_d2(Q.d, Q.c, Q.b, clampf(ra, 0.f, 1.f));
        const float db = quad_d2(Q.d, Q.c, Q.b, clampf(rb, 0.f, 1.f));
        Q.use_b = db < da;
        traw = Q.use_b ? rb : ra;
        const double kap = 1e-8 + 5e-9 / z, mu = p + 5e-9 / z, dkap = 5e-9 / (2.0 * z * z * z);
        Q.c1 = -3e-8f;
        Q.c0 = (float)(-q * kap / mu);
        Q.ep = (float)(-q * (dkap * mu - kap * (1.0 + dkap)) / (mu * mu));
        Q.eq = (float)(-kap / mu);
    }
    Q.root = (Q.one || (Q.use_b && !Q.clamped)) && traw > 0.f && traw < 1.f;
    float t = clampf(traw, 0.f, 1.f);
    if (Q.root) {
        const V3 v = Q.d + (Q.c + Q.b * t) * t;
        const V3 tg = Q.a + Q.b * t;
        const float F = dot(v, tg) * Q.kk + Q.c1 * (t + Q.kx) + Q.c0;
        const float Fp = (2.f * dot(tg, tg) + dot(v, Q.b)) * Q.kk + Q.c1;
        if (Fp > 0.f) t -= F / Fp;
        if (!(t > 0.f && t < 1.f)) {  // the root lies past an end: the pick clamps there
            Q.root = false;
            traw = t;
            t = clampf(t, 0.f, 1.f);
        } else {
            traw = t;
        }
    }
    Q.traw = traw;
    Q.t = t;
    Q.dist2 = quad_d2(Q.d, Q.c, Q.b, t);
    return Q;
}

__device__ __forceinline__ float eval_quad_bezier(V3 pos, const float* prm) {
    QuadSolve Q = quad_solve(pos, prm);
    return sqrtf(Q.dist2 + 1e-8f) - lerpf(prm[9], prm[10], Q.t);
}

// ---------------------------------------------------------------------------
// cubic Bezier / centripetal Catmull-Rom tubes: polyline projection then the
// exact curve point, strokes_sdf.h:1228-1383
// ---------------------------------------------------------------------------
__device__ __forceinline__ void bern3(float t, float w[4]) {
    float u = 1.f - t;
    w[0] = u * u * u;
    w[1] = 3.f * u * u * t;
    w[2] = 3.f * u * t * t;
    w[3] = t * t * t;
}
// curve in power form P(t) = K3 t^3 + K2 t^2 + K1 t + K0 for Catmull-Rom
__device__ __forceinline__ void powb(float t, float w[4]) {
    float t2 = t * t;
    w[0] = t2 * t;
    w[1] = t2;
    w[2] = t;
    w[3] = 1.f;
}
template <bool POWER>
__device__ __forceinline__ V3 curve_point(float t, const V3 K[4]) {
    float w[4];
    if (POWER) powb(t, w); else bern3(t, w);
    return w[0] * K[0] + w[1] * K[1] + w[2] * K[2] + w[3] * K[3];
}
template <bool POWER>
__device__ __forceinline__ V3 curve_tangent(float t, const V3 K[4]) {
    if (POWER) return (3.f * t * t) * K[0] + (2.f * t) * K[1] + K[2];
    float u = 1.f - t;
    return 3.f * ((u * u) * (K[1] - K[0]) + (2.f * u * t) * (K[2] - K[1]) + (t * t) * (K[3] - K[2]));
}

struct CurvePick {
    float t;     // t_min
    int seg;     // selected segment (0-based)
    float s;     // clamped local projection parameter
    float sraw;  // unclamped projection parameter
    float num, den;
    V3 S0, AB;
};

template <bool POWER>
__device__ __forceinline__ CurvePick curve_pick(V3 pos, const V3 K[4]) {
    constexpr float step = 1.0f / CURVE_SEGMENTS;
    CurvePick P;
    float dist_min = 1e18f;  // squared: the reference's distance bound of 1e9
    P.t = 0.f; P.seg = 0; P.s = 0.f; P.sraw = 0.f; P.num = 0.f; P.den = 1.f;
    P.S0 = v3s(0.f); P.AB = v3s(0.f);
    float t_start = 0.f;
    V3 s0 = curve_point<POWER>(t_start, K);
#pragma unroll
    for (int i = 1; i <= CURVE_SEGMENTS; i++) {
        float t_end = i * step;
        V3 s1 = curve_point<POWER>(t_end, K);
        V3 ab = s1 - s0;
        float num = dot(pos - s0, ab);
        float den = dot(ab, ab);
        float sr = num * __builtin_amdgcn_rcpf(den);
        float s = clampf(sr, 0.f, 1.f);
        V3 pp = s0 + s * ab;
        // squared distances order the segments as the distances do (sqrt is monotonic)
        V3 dv = pos - pp;
        float dist = dot(dv, dv);
        if (dist < dist_min) {
            dist_min = dist;
            P.t = t_start * (1.f - s) + t_end * s;
            P.seg = i - 1; P.s = s; P.sraw = sr; P.num = num; P.den = den; P.S0 = s0; P.AB = ab;
        }
        t_start = t_end;
        s0 = s1;
    }
    return P;
}

__device__ __forceinline__ void cubic_ctrl(const float* prm, V3 K[4]) {
    K[0] = ld3(prm); K[1] = ld3(prm + 3); K[2] = ld3(prm + 6); K[3] = ld3(prm + 9);
}

struct CatmullCoef {
    V3 p0, p1, p2, p3;
    float L01, L12, L23, t01, t12, t23;
    V3 m1, m2;
    V3 K[4];  // A, B, C(=m1), D(=p1)
};

__device__ __forceinline__ CatmullCoef catmull_coef(const float* prm) {
    CatmullCoef c;
    c.p0 = ld3(prm); c.p1 = ld3(prm + 3); c.p2 = ld3(prm + 6); c.p3 = ld3(prm + 9);
    c.L01 = safe_len(c.p0 - c.p1);
    c.L12 = safe_len(c.p1 - c.p2);
    c.L23 = safe_len(c.p2 - c.p3);
    c.t01 = sqrtf(c.L01 + 1e-8f);
    c.t12 = sqrtf(c.L12 + 1e-8f);
    c.t23 = sqrtf(c.L23 + 1e-8f);
    c.m1 = c.p2 - c.p1 + c.t12 * ((c.p1 - c.p0) * (1.f / c.t01) - (c.p2 - c.p0) * (1.f / (c.t01 + c.t12)));
    c.m2 = c.p2 - c.p1 + c.t12 * ((c.p3 - c.p2) * (1.f / c.t23) - (c.p3 - c.p1) * (1.f / (c.t12 + c.t23)));
    c.K[0] = 2.f * (c.p1 - c.p2) + c.m1 + c.m2;
    c.K[1] = 3.f * (c.p2 - c.p1) - 2.f * c.m1 - c.m2;
    c.K[2] = c.m1;
    c.K[3] = c.p1;
    return c;
}

template <bool POWER>
__device__ __forceinline__ float eval_curve(V3 pos, const V3 K[4], float r1, float r2) {
    CurvePick P = curve_pick<POWER>(pos, K);
    V3 Pt = curve_point<POWER>(P.t, K);
    return safe_len(Pt - pos) - lerpf(r1, r2, P.t);
}

// ---------------------------------------------------------------------------
// runtime dispatch: forward value
// ---------------------------------------------------------------------------
__device__ __forceinline__ float sdf_eval(int base, V3 p, const float* prm) {
    switch (base) {
        case SPHERE: return eval_sphere(p);
        case CUBE: return eval_cube(p);
        case ROUND_CUBE: return eval_round_cube(p, prm);
        case CAPPED_TORUS: return eval_capped_torus(p, prm);
        case CAPSULE: return eval_capsule(p, prm);
        case LINE: return eval_line(p, prm);
        case TRIPRISM: return eval_triprism(p, prm);
        case OCTAHEDRON: return eval_octahedron(p);
        case TETRAHEDRON: return eval_tetrahedron(p);
        case QUAD_BEZIER: return eval_quad_bezier(p, prm);
        case CUBIC_BEZIER: {
            V3 K[4];
            cubic_ctrl(prm, K);
            return eval_curve<false>(p, K, prm[12], prm[13]);
        }
        default: {
            CatmullCoef c = catmull_coef(prm);
            return eval_curve<true>(p, c.K, prm[12], prm[13]);
        }
    }
}

// ---------------------------------------------------------------------------
// exact adjoints of the forward values
// ---------------------------------------------------------------------------
__device__ __forceinline__ V3 face_normal(V3 q, V3 p) {
    // tie-breaking of the reference's inside-cube branch (strokes_sdf.h:284-287)
    float gx = (q.x >= q.y && q.x >= q.z) ? sgn(p.x) : 0.f;
    float gy = (q.x < q.y && q.y >= q.z) ? sgn(p.y) : 0.f;
    float gz = (q.x < q.z && q.y < q.z) ? sgn(p.z) : 0.f;
    return v3(gx, gy, gz);
}

// quadratic Bezier: exact derivative of eval_quad_bezier.  For refined roots
// dt/dtheta = -(dF/dtheta) / F'(t) (implicit function theorem on the cubic the
// closed form solves); for the three-root branch's first candidate the
// reverse-mode derivative of the closed form (acos branch, its regularisers).
__device__ __forceinline__ V3 grad_ex_quad_solved(V3 pos, const float* prm, const QuadSolve& Q, float g, float* gp) {
    const float t = Q.t;
    const float r1 = prm[9], r2 = prm[10];
    const V3 v = Q.d + (Q.c + Q.b * t) * t;  // P(t) - pos
    const V3 tg = Q.a + Q.b * t;             // P'(t) / 2
    const float dist = sqrtf(dot(v, v) + 1e-8f);
    const V3 vn = v * (1.f / dist);
    const float G = 2.f * dot(v, tg) / dist - (r2 - r1);  // dSDF/dt at fixed geometry
    const float kk = Q.kk;
    // direct dependence of v = d + 2 a t + b t^2 at fixed t
    V3 ga = vn * (2.f * t), gb = vn * (t * t), gd = vn;
    float gkx = 0.f, gky = 0.f, gkz = 0.f;
    if (Q.root) {
        const float gg = dot(v, tg);
        const float Fp = (2.f * dot(tg, tg) + dot(v, Q.b)) * kk + Q.c1;
        const float lam = Fp > 0.f ? -G / Fp : 0.f;
        // g / |b|^2 part of F (kk = 1 / |b|^2)
        ga = ga + (2.f * t * tg + v) * (lam * kk);
        gb = gb + (t * t * tg + t * v) * (lam * kk) - Q.b * (2.f * lam * gg * kk * kk);
        gd = gd + tg * (lam * kk);
        // regulariser part C(kx, p, q) = c1 (t + kx) + c0
        const float Cp = lam * (Q.one ? Q.ep * (t + Q.kx) : Q.ep), Cq = lam * Q.eq;
        gkx = lam * Q.c1 + Cq * (6.f * Q.kx * Q.kx - 3.f * Q.ky) - 2.f * Q.kx * Cp;
        gky = -3.f * Q.kx * Cq + Cp;
        gkz = Cq;
    } else if (!Q.one) {
        const float sg = Q.use_b ? -1.f : 1.f;
        const float gt = (Q.traw > 0.f && Q.traw < 1.f) ? G : 0.f;
        float gm = gt * sg * Q.z, gn = gt * sg * Q.z, gz = gt * sg * (Q.m + Q.n);
        float gv = -sinf(Q.v) * gm + 1.732050808f * cosf(Q.v) * gn;
        // d acos(arg) = -d arg / sqrt(1 - arg^2); zero where the argument is clamped
        float garg = (!Q.clamped && Q.sq > 0.f) ? -gv / (3.f * Q.sq) : 0.f;
        float gq = garg / Q.den;
        float gden = -garg * Q.q / (Q.den * Q.den);
        float gpp = 2.f * Q.z * gden;
        gz += 2.f * Q.p * gden;
        gpp += -gz / (2.f * Q.z);
        // q = 2kx^3 - 3 kx ky + kz, p = ky - kx^2
        gkx = -gt + gq * (6.f * Q.kx * Q.kx - 3.f * Q.ky) - 2.f * Q.kx * gpp;
        gky = -3.f * Q.kx * gq + gpp;
        gkz = gq;
    }
    // kx = kk ab, ky = kk (2 aa + db)/3, kz = kk da, kk = 1/bb
    const float ab = dot(Q.a, Q.b), aa = dot(Q.a, Q.a), db = dot(Q.d, Q.b), da = dot(Q.d, Q.a);
    const float gkk = gkx * ab + gky * (2.f * aa + db) / 3.f + gkz * da;
    const float gab = gkx * kk, gaa = gky * kk * (2.f / 3.f), gdb = gky * kk * (1.f / 3.f), gda = gkz * kk;
    const float gbb = -gkk * kk * kk;
    ga = ga + gab * Q.b + (2.f * gaa) * Q.a + gda * Q.d;
    gb = gb + gab * Q.a + gdb * Q.d + (2.f * gbb) * Q.b;
    gd = gd + gdb * Q.b + gda * Q.a;
    // a = B - A, b = A - 2B + C, d = A - pos
    add3(gp + 0, (-ga + gb + gd) * g);
    add3(gp + 3, (ga - 2.f * gb) * g);
    add3(gp + 6, gb * g);
    gp[9] += -(1.f - t) * g;
    gp[10] += -t * g;
    return -gd;
}

__device__ __forceinline__ V3 grad_ex_quad(V3 pos, const float* prm, float g, float* gp) {
    const QuadSolve Q = quad_solve(pos, prm);
    return grad_ex_quad_solved(pos, prm, Q, g, gp);
}

// cubic Bezier / Catmull-Rom: derivative through the polyline parameter.
// Writes dSDF/dK (K = control points for Bezier, power coefficients for
// Catmull-Rom) into gK and returns dSDF/dpos.
template <bool POWER>
__device__ __forceinline__ V3 grad_ex_curve_picked(V3 pos, const V3 K[4], float r1, float r2, const CurvePick& P,
                                                  V3 gK[4], float& dr1, float& dr2) {
    constexpr float step = 1.0f / CURVE_SEGMENTS;
    float t = P.t;
    V3 Pt = curve_point<POWER>(t, K);
    V3 v = Pt - pos;
    float dist = safe_len(v);
    V3 vn = dist > 0.f ? v * (1.f / dist) : v3s(0.f);
    V3 Pp = curve_tangent<POWER>(t, K);
    float G = dot(vn, Pp) - (r2 - r1);
    float w[4];
    if (POWER) powb(t, w); else bern3(t, w);
#pragma unroll
    for (int j = 0; j < 4; j++) gK[j] = vn * w[j];
    V3 gpos = -vn;
    if (P.sraw > 0.f && P.sraw < 1.f) {
        // s = num/den, num = (pos - S0)·AB, den = AB·AB; t = t_s + s·Δt
        float id = 1.f / P.den;
        V3 rel = pos - P.S0;
        V3 ds_dS0 = (-P.AB - rel) * id + P.AB * (2.f * P.num * id * id);
        V3 ds_dS1 = rel * id - P.AB * (2.f * P.num * id * id);
        float ts = P.seg * step, te = (P.seg + 1) * step;
        float Gd = G * (te - ts);
        float ws[4], we[4];
        if (POWER) { powb(ts, ws); powb(te, we); } else { bern3(ts, ws); bern3(te, we); }
#pragma unroll
        for (int j = 0; j < 4; j++) gK[j] = gK[j] + Gd * (ws[j] * ds_dS0 + we[j] * ds_dS1);
        gpos = gpos + Gd * (P.AB * id);
    }
    dr1 = -(1.f - t);
    dr2 = -t;
    return gpos;
}

template <bool POWER>
__device__ __forceinline__ V3 grad_ex_curve_K(V3 pos, const V3 K[4], float r1, float r2, V3 gK[4], float& dr1,
                                             float& dr2) {
    const CurvePick P = curve_pick<POWER>(pos, K);
    return grad_ex_curve_picked<POWER>(pos, K, r1, r2, P, gK, dr1, dr2);
}


// Catmull-Rom coefficient adjoint: dL/dK -> dL/d(p0..p3)
__device__ __forceinline__ void catmull_coef_adjoint(const CatmullCoef& c, const V3 gK[4], V3 gp[4]) {
    V3 gA = gK[0], gB = gK[1], gC = gK[2], gD = gK[3];
    V3 gm1 = gC + gA - 2.f * gB;
    V3 gm2 = gA - gB;
    V3 g0 = v3s(0.f), g1 = gD + 2.f * gA - 3.f * gB, g2 = -2.f * gA + 3.f * gB, g3 = v3s(0.f);
    float gt01 = 0.f, gt12 = 0.f, gt23 = 0.f;
    // m1 = p2 - p1 + t12 * (u1 - v1), u1 = (p1 - p0)/t01, v1 = (p2 - p0)/(t01 + t12)
    {
        float s1 = c.t01 + c.t12;
        V3 u1 = (c.p1 - c.p0) * (1.f / c.t01);
        V3 v1 = (c.p2 - c.p0) * (1.f / s1);
        g2 = g2 + gm1;
        g1 = g1 - gm1;
        gt12 += dot(gm1, u1 - v1);
        V3 gu = gm1 * c.t12, gv = gm1 * (-c.t12);
        g1 = g1 + gu * (1.f / c.t01);
        g0 = g0 - gu * (1.f / c.t01);
        gt01 += -dot(gu, u1) / c.t01;
        g2 = g2 + gv * (1.f / s1);
        g0 = g0 - gv * (1.f / s1);
        float gs = -dot(gv, v1) / s1;
        gt01 += gs;
        gt12 += gs;
    }
    // m2 = p2 - p1 + t12 * (u2 - v2), u2 = (p3 - p2)/t23, v2 = (p3 - p1)/(t12 + t23)
    {
        float s2 = c.t12 + c.t23;
        V3 u2 = (c.p3 - c.p2) * (1.f / c.t23);
        V3 v2 = (c.p3 - c.p1) * (1.f / s2);
        g2 = g2 + gm2;
        g1 = g1 - gm2;
        gt12 += dot(gm2, u2 - v2);
        V3 gu = gm2 * c.t12, gv = gm2 * (-c.t12);
        g3 = g3 + gu * (1.f / c.t23);
        g2 = g2 - gu * (1.f / c.t23);
        gt23 += -dot(gu, u2) / c.t23;
        g3 = g3 + gv * (1.f / s2);
        g1 = g1 - gv * (1.f / s2);
        float gs = -dot(gv, v2) / s2;
        gt12 += gs;
        gt23 += gs;
    }
    // tij = sqrt(Lij + 1e-8), Lij = |pi - pj|
    auto back = [](float gt, float tij, float L, V3 pi, V3 pj, V3& gi, V3& gj) {
        if (L > 0.f) {
            float gL = gt * 0.5f / tij;
            V3 dv = (pi - pj) * (gL / L);
            gi = gi + dv;
            gj = gj - dv;
        }
    };
    back(gt01, c.t01, c.L01, c.p0, c.p1, g0, g1);
    back(gt12, c.t12, c.L12, c.p1, c.p2, g1, g2);
    back(gt23, c.t23, c.L23, c.p2, c.p3, g2, g3);
    gp[0] = g0; gp[1] = g1; gp[2] = g2; gp[3] = g3;
}

// Exact SDF value of a curve base together with its adjoint from one
// closest-point solve: dsdf(sdf) maps the value to dL/dSDF (the caller's
// transfer derivative), the parameter gradients are accumulated into gp.
template <typename DSDF>
__device__ __forceinline__ V3 curve_value_grad(int base, V3 pos, const float* prm, DSDF dsdf, float* gp, float& sdf,
                                               float& g) {
    if (base == QUAD_BEZIER) {
        const QuadSolve Q = quad_solve(pos, prm);
        sdf = sqrtf(Q.dist2 + 1e-8f) - lerpf(prm[9], prm[10], Q.t);
        g = dsdf(sdf);
        return grad_ex_quad_solved(pos, prm, Q, g, gp);
    }
    V3 gK[4];
    float dr1, dr2;
    V3 gpos;
    if (base == CUBIC_BEZIER) {
        V3 K[4];
        cubic_ctrl(prm, K);
        const CurvePick P = curve_pick<false>(pos, K);
        sdf = safe_len(curve_point<false>(P.t, K) - pos) - lerpf(prm[12], prm[13], P.t);
        g = dsdf(sdf);
        gpos = grad_ex_curve_picked<false>(pos, K, prm[12], prm[13], P, gK, dr1, dr2);
    } else {
        const CatmullCoef cc = catmull_coef(prm);
        const CurvePick P = curve_pick<true>(pos, cc.K);
        sdf = safe_len(curve_point<true>(P.t, cc.K) - pos) - lerpf(prm[12], prm[13], P.t);
        g = dsdf(sdf);
        V3 gK2[4];
        gpos = grad_ex_curve_picked<true>(pos, cc.K, prm[12], prm[13], P, gK2, dr1, dr2);
        catmull_coef_adjoint(cc, gK2, gK);
    }
#pragma unroll
    for (int j = 0; j < 4; j++) add3(gp + 3 * j, gK[j] * g);
    gp[12] += dr1 * g;
    gp[13] += dr2 * g;
    return gpos;
}

__device__ __forceinline__ V3 grad_ex(int base, V3 p, const float* prm, float g, float* gp) {
    switch (base) {
        case SPHERE: {
            float inv = 1.f / sqrtf(p.x * p.x + p.y * p.y + p.z * p.z + 1e-8f);
            return p * inv;
        }
        case CUBE: {
            V3 q = vabs(p) - v3s(1.f);
            V3 qp = vmax0(q);
            float l2 = dot(qp, qp);
            if (l2 > 0.f) {
                float inv = 1.f / sqrtf(l2);
                return mul(qp * inv, vsgn(p));
            }
            return face_normal(q, p);
        }
        case ROUND_CUBE: {
            V3 d = vabs(p) - v3s(1.f);
            V3 dp = vmax0(d);
            float inv = 1.f / sqrtf(dp.x * dp.x + dp.y * dp.y + dp.z * dp.z + 1e-8f);
            V3 gr = mul(dp * inv, vsgn(p));
            float mx = fmaxf(d.x, fmaxf(d.y, d.z));
            if (mx <= 0.f) gr = gr + face_normal(d, p);
            gp[0] += -g;
            return gr;
        }
        case CAPPED_TORUS: {
            float ax = fabsf(p.x);
            float s = sinf(prm[0]), c = cosf(prm[0]);
            bool cr = (ax * c - p.y * s) > 0.f;
            float ln = sqrtf(ax * ax + p.y * p.y + 1e-8f);
            float k = cr ? ax * s + p.y * c : ln;
            float inv = 1.f / sqrtf(ax * ax + p.y * p.y + p.z * p.z + 1.f - 2.f * k + 1e-8f);
            float dkx = cr ? s * sgn(p.x) : p.x / ln;
            float dky = cr ? c : p.y / ln;
            if (cr) gp[0] += -g * (ax * c - p.y * s) * inv;
            gp[1] += -g;
            return v3((p.x - dkx) * inv, (p.y - dky) * inv, p.z * inv);
        }
        case CAPSULE: {
            float h = prm[0];
            float y = p.y - clampf(p.y, -h, h);
            float inv = 1.f / sqrtf(p.x * p.x + y * y + p.z * p.z + 1e-8f);
            bool inside = (-h < p.y) && (p.y < h);
            if (!inside) gp[0] += g * y * inv * (p.y > 0.f ? -1.f : 1.f);
            return v3(p.x * inv, inside ? 0.f : y * inv, p.z * inv);
        }
        case LINE: {
            float h = prm[0], rd = prm[1];
            float traw = (p.y + h) / (2.f * h);
            float t = clampf(traw, 0.f, 1.f);
            float y = p.y - clampf(p.y, -h, h);
            float inv = 1.f / sqrtf(p.x * p.x + y * y + p.z * p.z + 1e-8f);
            bool inside = (-h < p.y) && (p.y < h);
            bool tin = (traw > 0.f) && (traw < 1.f);
            float gy = inside ? 0.f : y * inv;
            float gh = inside ? 0.f : y * inv * (p.y > 0.f ? -1.f : 1.f);
            if (tin) {
                gy += -2.f * rd / (2.f * h);
                gh += -2.f * rd * (-p.y / (2.f * h * h));
            }
            gp[0] += g * gh;
            gp[1] += g * (1.f - 2.f * t);
            return v3(p.x * inv, gy, p.z * inv);
        }
        case TRIPRISM: {
            float qx = fabsf(p.x), qy = fabsf(p.y);
            float e1 = qx * TRI_C + p.z * 0.5f;
            float inner = fmaxf(e1, -p.z) - 0.5f;
            if (qy - prm[0] > inner) {
                gp[0] += -g;
                return v3(0.f, sgn(p.y), 0.f);
            }
            if (e1 > -p.z) return v3(TRI_C * sgn(p.x), 0.f, 0.5f);
            return v3(0.f, 0.f, -1.f);
        }
        case OCTAHEDRON:
            return vsgn(p) * OCTA_C;
        case TETRAHEDRON: {
            float a = p.x + p.y, c = p.x - p.y;
            if (fabsf(a) - p.z > fabsf(c) + p.z) return v3(sgn(a) * OCTA_C, sgn(a) * OCTA_C, -OCTA_C);
            return v3(sgn(c) * OCTA_C, -sgn(c) * OCTA_C, OCTA_C);
        }
        case QUAD_BEZIER:
            return grad_ex_quad(p, prm, g, gp);
        case CUBIC_BEZIER: {
            V3 K[4], gK[4];
            cubic_ctrl(prm, K);
            float dr1, dr2;
            V3 gpos = grad_ex_curve_K<false>(p, K, prm[12], prm[13], gK, dr1, dr2);
#pragma unroll
            for (int j = 0; j < 4; j++) add3(gp + 3 * j, gK[j] * g);
            gp[12] += dr1 * g;
            gp[13] += dr2 * g;
            return gpos;
        }
        default: {
            CatmullCoef cc = catmull_coef(prm);
            V3 gK[4], gq[4];
            float dr1, dr2;
            V3 gpos = grad_ex_curve_K<true>(p, cc.K, prm[12], prm[13], gK, dr1, dr2);
            catmull_coef_adjoint(cc, gK, gq);
#pragma unroll
            for (int j = 0; j < 4; j++) add3(gp + 3 * j, gq[j] * g);
            gp[12] += dr1 * g;
            gp[13] += dr2 * g;
            return gpos;
        }
    }
}

// ---------------------------------------------------------------------------
// the reference's hand-written adjoints (strokes_sdf.h grad_sdf members)
// ---------------------------------------------------------------------------
__device__ V3 grad_ref_quad(V3 pos, const float* prm, float g, float* gp);

__device__ __forceinline__ V3 grad_ref(int base, V3 p, const float* prm, float g, float* gp) {
    switch (base) {
        case SPHERE: {
            float inv = rsqrtf(p.x * p.x + p.y * p.y + p.z * p.z + 1e-8f);
            return p * inv;
        }
        case CUBE: {  // strokes_sdf.h:279-299: outside uses 2·max(q,0)
            V3 q = vabs(p) - v3s(1.f);
            if (fmaxf(q.x, fmaxf(q.y, q.z)) <= 0.f) return face_normal(q, p);
            V3 gl = vmax0(q) * 2.f;
            V3 gq = v3(q.x > 0.f ? gl.x : 0.f, q.y > 0.f ? gl.y : 0.f, q.z > 0.f ? gl.z : 0.f);
            return v3(p.x < 0.f ? -gq.x : gq.x, p.y < 0.f ? -gq.y : gq.y, p.z < 0.f ? -gq.z : gq.z);
        }
        case ROUND_CUBE: {  // strokes_sdf.h:327-355
            gp[0] += -g;
            V3 d = vabs(p) - v3s(1.f);
            V3 dp = vmax0(d);
            float sq = dp.x * dp.x + dp.y * dp.y + dp.z * dp.z;
            float n = sqrtf(sq + 1e-8f);
            float inv = rsqrtf(sq + 1e-8f);
            V3 gdp = n > 0.f ? dp * inv : v3s(0.f);
            V3 sg = vsgn(p);
            V3 g1 = mul(gdp, sg);
            float mn = fminf(fmaxf(d.x, fmaxf(d.y, d.z)), 0.f);
            float gm = mn < 0.f ? 1.f : 0.f;
            float gx = gm * ((d.x >= d.y && d.x >= d.z) ? 1.f : 0.f) * sg.x;
            float gy = gm * ((d.x < d.y && d.y >= d.z) ? 1.f : 0.f) * sg.y;
            float gz = gm * ((d.x < d.z && d.y < d.z) ? 1.f : 0.f) * sg.z;
            return v3(g1.x + gx, g1.y + gy, g1.z + gz);
        }
        case CAPPED_TORUS: {  // strokes_sdf.h:382-420
            float ax = fabsf(p.x);
            float s = sinf(prm[0]), c = cosf(prm[0]);
            float k = (ax * c - p.y * s) > 0.f ? ax * s + p.y * c : safe_len(v3(ax, p.y, 0.f));
            float inv = rsqrtf(ax * ax + p.y * p.y + p.z * p.z + 1.f - 2.f * k + 1e-8f);
            V3 r;
            if (k > 0.f) {
                gp[0] += -g * (ax * c + p.y * s);
                r = v3((p.x + (-sgn(p.x) * s)) * inv, (p.y + (-c)) * inv, p.z * inv);
            } else {
                r = v3((p.x + p.x) * inv, (p.y + p.y) * inv, p.z * inv);
            }
            gp[1] += -g;
            return r;
        }
        case CAPSULE: {  // strokes_sdf.h:442-454
            float h = prm[0];
            float y = p.y - clampf(p.y, -h, h);
            float inv = rsqrtf(p.x * p.x + y * y + p.z * p.z + 1e-8f);
            V3 r = v3(p.x * inv, y * inv, p.z * inv);
            if (-h <= p.y && p.y <= h) r.y = 0.f;
            else gp[0] += g * r.y * (p.y > 0.f ? -1.f : 1.f);
            return r;
        }
        case LINE: {  // strokes_sdf.h:486-509
            float h = prm[0], rd = prm[1];
            float t = clampf((p.y + h) / (2.f * h), 0.f, 1.f);
            float y = p.y - clampf(p.y, -h, h);
            float inv = rsqrtf(p.x * p.x + y * y + p.z * p.z + 1e-8f);
            V3 r = v3(p.x * inv, y * inv, p.z * inv);
            float g0;
            if (-h <= p.y && p.y <= h) {
                r.y = -rd / h;
                g0 = p.y * rd / (h * h);
            } else {
                g0 = r.y * (p.y > 0.f ? -1.f : 1.f);
            }
            gp[0] += g * g0;
            gp[1] += g * (1.f - 2.f * t);
            return r;
        }
        case TRIPRISM: {  // strokes_sdf.h:537-572
            float qx = fabsf(p.x), qy = fabsf(p.y);
            float e1 = qx * TRI_C + p.z * 0.5f;
            if (e1 > -p.z) {
                if (qy - prm[0] > e1 - 0.5f) { gp[0] += -g; return v3(0.f, sgn(p.y), 0.f); }
                return v3(TRI_C * sgn(p.x), 0.f, 0.5f);
            }
            if (qy - prm[0] > -p.z - 0.5f) { gp[0] += -g; return v3(0.f, sgn(p.y), 0.f); }
            return v3(0.f, 0.f, -1.f);
        }
        case OCTAHEDRON:
            return vsgn(p) * OCTA_C;
        case TETRAHEDRON: {
            float a = p.x + p.y, c = p.x - p.y;
            if (fabsf(a) - p.z > fabsf(c) + p.z) return v3(sgn(a) * OCTA_C, sgn(a) * OCTA_C, -OCTA_C);
            return v3(sgn(c) * OCTA_C, -sgn(c) * OCTA_C, OCTA_C);
        }
        case QUAD_BEZIER:
            return grad_ref_quad(p, prm, g, gp);
        case CUBIC_BEZIER: {  // strokes_sdf.h:1285-1318 (squared-distance gradient, t held fixed)
            V3 K[4];
            cubic_ctrl(prm, K);
            CurvePick P = curve_pick<false>(p, K);
            float t = P.t, u = 1.f - t;
            V3 v = curve_point<false>(t, K) - p;
            V3 gv = v * (g * 2.f);
            float w[4];
            bern3(t, w);
#pragma unroll
            for (int j = 0; j < 4; j++) add3(gp + 3 * j, gv * w[j]);
            gp[12] += -g * u;
            gp[13] += -g * t;
            return -gv;
        }
        default: {  // Catmull-Rom, strokes_sdf.h:1385-1445
            CatmullCoef c = catmull_coef(prm);
            CurvePick P = curve_pick<true>(p, c.K);
            float t = P.t;
            V3 v = curve_point<true>(t, c.K) - p;
            V3 gv = v * (g * 2.f);
            V3 p0p1 = c.p0 - c.p1, p1p2 = c.p1 - c.p2, p2p3 = c.p2 - c.p3;
            float t01 = c.t01, t12 = c.t12, t23 = c.t23;
            V3 gA = gv * (t * t * t), gB = gv * (t * t), gC = gv * t, gD = gv;
            V3 gm1 = gA - gB * 2.f + gC;
            V3 gm2 = gA - gB;
            float dm1_dp0 = t12 / (t01 + t12) - t12 / t01;
            float dm1_dp1 = -1.f + t12 / t01;
            float dm1_dp2 = 1.f - t12 / (t01 + t12);
            float dm2_dp1 = -1.f + t12 / (t12 + t23);
            float dm2_dp2 = 1.f - t12 / t23;
            float dm2_dp3 = t12 / t23 - t12 / (t12 + t23);
            float s01 = (t01 + t12) * (t01 + t12), s23 = (t12 + t23) * (t12 + t23);
            V3 dm1_dt01 = t12 * ((c.p2 - c.p0) * (1.f / s01) - (c.p1 - c.p0) * (1.f / (t01 * t01)));
            V3 dm1_dt12 = (c.p1 - c.p0) * (1.f / t01) + t12 * (c.p2 - c.p0) * (1.f / s01) - (c.p2 - c.p0) * (1.f / (t01 + t12));
            V3 dm2_dt12 = (c.p2 - c.p1) * (1.f / t23) + t12 * (c.p3 - c.p1) * (1.f / s23) - (c.p3 - c.p1) * (1.f / (t12 + t23));
            V3 dm2_dt23 = t12 * ((c.p2 - c.p3) * (1.f / (t23 * t23)) + (c.p3 - c.p1) * (1.f / s23));
            V3 dt01 = p0p1 * (1.f / (2.f * powf(dot(p0p1, p0p1) + 1e-8f, 0.75f)));
            V3 dt12 = p1p2 * (1.f / (2.f * powf(dot(p1p2, p1p2) + 1e-8f, 0.75f)));
            V3 dt23 = p2p3 * (1.f / (2.f * powf(dot(p2p3, p2p3) + 1e-8f, 0.75f)));
            V3 g0 = mul(gm1, v3s(dm1_dp0) + mul(dm1_dt01, dt01));
            V3 g1 = gA * 2.f + gB * -3.f + gD + mul(gm1, v3s(dm1_dp1) - mul(dm1_dt01, dt01) + mul(dm1_dt12, dt12)) +
                    mul(gm2, v3s(dm2_dp1) + mul(dm2_dt12, dt12));
            V3 g2 = gA * -2.f + gB * 3.f + mul(gm1, v3s(dm1_dp2) - mul(dm1_dt12, dt12)) +
                    mul(gm2, v3s(dm2_dp2) - mul(dm2_dt12, dt12) + mul(dm2_dt23, dt23));
            V3 g3 = mul(gm2, v3s(dm2_dp3) - mul(dm2_dt23, dt23));
            add3(gp + 0, g0); add3(gp + 3, g1); add3(gp + 6, g2); add3(gp + 9, g3);
            gp[12] += -g * (1.f - t);
            gp[13] += -g * t;
            return -gv;
        }
    }
}

// Reference quadratic Bezier adjoint, strokes_sdf.h:708-1172, restated with its
// comma-operator partials (x/6 in the position path, x/3 in the parameter
// path) and the (t + 1) radius factor of the three-root / first-candidate branch.
__device__ inline V3 grad_ref_quad(V3 pos, const float* prm, float g, float* gp) {
    V3 A = ld3(prm), B = ld3(prm + 3), C = ld3(prm + 6);
    float r1 = prm[9], r2 = prm[10];
    V3 a = B - A, b = A - 2.f * B + C, c = a * 2.f, d = A - pos;
    float kk = 1.f / dot(b, b);
    float kx = kk * dot(a, b);
    float ky = kk * (2.f * dot(a, a) + dot(d, b)) / 3.f;
    float kz = kk * dot(d, a);
    float p = ky - kx * kx;
    float q = kx * (2.f * kx * kx - 3.f * ky) + kz;
    float s = q * q + 4.f * p * p * p;
    float b2 = dot(b, b);
    float gkk_b2 = -1.f / (b2 * b2);
    float ab = dot(a, b), aa = dot(a, a), db = dot(d, b), da = dot(d, a);
    // per-parameter partials (parameter j -> unit direction e_{j%3} of A, B or C)
    float g_kx[9], g_p[9], g_q[9];
    const float wa[3] = {-1.f, 1.f, 0.f}, wb[3] = {1.f, -2.f, 1.f}, wc[3] = {-2.f, 2.f, 0.f}, wd[3] = {1.f, 0.f, 0.f};
    const float bc[3] = {b.x, b.y, b.z}, ac[3] = {a.x, a.y, a.z}, dc[3] = {d.x, d.y, d.z};
#pragma unroll
    for (int j = 0; j < 9; j++) {
        int pt = j / 3, k = j % 3;
        float gb2 = 2.f * bc[k] * wb[pt];
        float gkk = gkk_b2 * gb2;
        float gab = bc[k] * wa[pt] + ac[k] * wb[pt];
        g_kx[j] = kk * gab + ab * gkk;
        float gaa = 2.f * ac[k] * wa[pt];
        float gdb = bc[k] * wd[pt] + dc[k] * wb[pt];
        float gky = 2.f / 3.f * kk * gaa + 1.f / 3.f * kk * gdb + 1.f / 3.f * (2 * aa + db) * gkk;
        float gda = dc[k] * wa[pt] + ac[k] * wd[pt];
        float gkz = kk * gda + da * gkk;
        g_p[j] = gky - 2.f * kx * g_kx[j];
        g_q[j] = 6.f * kx * kx * g_kx[j] - 3.f * ky * g_kx[j] - 3.f * kx * gky + gkz;
    }
    V3 g_ky_xyz = (1.f / 3.f * kk) * (-b);
    V3 g_kz_xyz = kk * (-a);
    V3 g_p_xyz = g_ky_xyz;
    V3 g_q_xyz = (-3.f * kx) * g_ky_xyz + g_kz_xyz;

    float tsel, res;
    V3 mv;
    int cat;           // 1, 2 or 3
    bool bug = false;  // (t + 1) radius factor
    float zz = 0.f, vv = 0.f, mm = 0.f, nn = 0.f, x0 = 0.f, x1 = 0.f;
    if (s >= 0.f) {
        float h = sqrtf(s + 1e-8f);
        x0 = (h - q) * 0.5f;
        x1 = -(h + q) * 0.5f;
        tsel = clampf((cbrt_signed(x0) + cbrt_signed(x1)) - kx, 0.f, 1.f);
        mv = d + (c + b * tsel) * tsel;
        res = dot(mv, mv);
        cat = (tsel == 0.f || tsel == 1.f) ? 3 : 1;
    } else {
        zz = sqrtf(-p + 1e-8f);
        vv = acosf(clampf(q / (p * zz * 2.f + 1e-8f), -1.f, 1.f)) / 3.f;
        mm = cosf(vv);
        nn = sinf(vv) * 1.732050808f;
        float ta = clampf((mm + nn) * zz - kx, 0.f, 1.f);
        float tb = clampf((-mm - nn) * zz - kx, 0.f, 1.f);
        V3 ma = d + (c + b * ta) * ta, mb = d + (c + b * tb) * tb;
        float disa = dot(ma, ma), disb = dot(mb, mb);
        if (disb < disa) { tsel = tb; mv = mb; res = disb; }
        else { tsel = ta; mv = ma; res = disa; bug = !(ta == 0.f || ta == 1.f); }
        cat = (tsel == 0.f || tsel == 1.f) ? 3 : 2;
    }
    float t = tsel;
    V3 gret;
    if (cat == 3) {
        float rs = 0.5f * rsqrtf(res + 1e-8f);
        gret = d * (-1.f * rsqrtf(res + 1e-8f));
        V3 dcb = d + c + b;
        const float dd[3] = {d.x, d.y, d.z}, ee[3] = {dcb.x, dcb.y, dcb.z};
#pragma unroll
        for (int j = 0; j < 9; j++) {
            int pt = j / 3, k = j % 3;
            float v0 = 2.f * dd[k] * wd[pt];
            float v1 = 2.f * ee[k] * (wd[pt] + wc[pt] + wb[pt]);
            gp[j] += g * ((t == 0.f ? v0 : v1) * rs);
        }
    } else {
        float rs = 0.5f * rsqrtf(res + 1e-8f);
        V3 gmt = c + 2.f * b * t;
        V3 gtx;
        float gt[9];
        if (cat == 1) {
            float gh_s = 0.5f * rsqrtf(s + 1e-8f);
            float gux_x = x0 / 6.f, guy_x = -1.f * x1 / 6.f;
            float gux_p = x0 / 3.f, guy_p = -1.f * x1 / 3.f;
            auto tg = [&](float gux, float guy, float gq, float gpp) {
                return (-gux * gq + gux * gh_s * (2.f * q * gq + 12.f * p * p * gpp)) +
                       (-guy * gq + guy * gh_s * (2.f * q * gq + 12.f * p * p * gpp));
            };
            gtx = v3(tg(gux_x, guy_x, g_q_xyz.x, g_p_xyz.x), tg(gux_x, guy_x, g_q_xyz.y, g_p_xyz.y),
                     tg(gux_x, guy_x, g_q_xyz.z, g_p_xyz.z));
#pragma unroll
            for (int j = 0; j < 9; j++) gt[j] = tg(gux_p, guy_p, g_q[j], g_p[j]) - g_kx[j];
        } else {
            float gtn = -1.f * zz, gtm = -1.f * zz, gtz = -mm - nn;
            float gzp = -0.5f * rsqrtf(-p + 1e-8f);
            float gnv = cosf(vv) * 1.732050808f, gmv = -sinf(vv);
            float den = p * zz * 2.f + 1e-8f;
            float rq = rsqrtf(1.f - q * q / (p * p * zz * zz * 4.f) + 1e-8f);
            float gvq = 1.f / den * (-1.f) * rq / 3.f;
            float gvp = 1.f / p * q / den * rq / 3.f;
            float gvz = 1.f / zz * q / den * rq / 3.f;
            V3 gvx = g_q_xyz * gvq + g_p_xyz * gvp + g_p_xyz * (gvz * gzp);
            gtx = (gvx * (gtn * gnv) + gvx * (gtm * gmv)) + g_p_xyz * (gtz * gzp);
#pragma unroll
            for (int j = 0; j < 9; j++) {
                float gvpi = gvq * g_q[j] + gvp * g_p[j] + gvz * gzp * g_p[j];
                gt[j] = (gtn * gnv * gvpi + gtm * gmv * gvpi) + gtz * gzp * g_p[j] - g_kx[j];
            }
        }
        // row vector 2m times (-I + outer(gtx, gmt))
        V3 tm = mv * 2.f;
        float sdot = dot(tm, gtx);
        V3 rx = v3(-tm.x + sdot * gmt.x, -tm.y + sdot * gmt.y, -tm.z + sdot * gmt.z) * rs;
        gret = rx - (r2 - r1) * gtx;
#pragma unroll
        for (int j = 0; j < 9; j++) {
            int pt = j / 3, k = j % 3;
            V3 e = v3(k == 0 ? 1.f : 0.f, k == 1 ? 1.f : 0.f, k == 2 ? 1.f : 0.f);
            V3 ghp = e * wd[pt] + c * gt[j] + e * (wc[pt] * t) + b * (2.f * t * gt[j]) + e * (wb[pt] * t * t);
            float val = dot(tm, ghp) * rs;
            gp[j] += g * (val - (r2 - r1) * gt[j]);
        }
    }
    gp[9] += -g * (bug ? (t + 1.f) : (-t + 1.f));
    gp[10] += -g * t;
    return gret;
}

// ---------------------------------------------------------------------------
// inv_scale_radius with per-axis scale: factor f = |∇ ⊙ s| / |∇| using the
// direction of the reference adjoint (strokes_forward.cu:73-80), and its
// exact derivative for the bases whose direction varies smoothly.
// ---------------------------------------------------------------------------
__device__ __forceinline__ V3 ref_dir(int base, V3 p, const float* prm) {
    float dummy[16];
    return grad_ref(base, p, prm, 0.f, dummy);
}
__device__ __forceinline__ float radius_factor(V3 gdir, V3 s) {
    V3 gs = mul(gdir, s);
    return safe_len(gs) / safe_len(gdir);
}
// dL/dp and dL/ds contributions from dL/df (exact mode).  The direction u(p)
// is differentiated for sphere (u ∝ p), cube / round cube outside
// (u ∝ max(|p|-1,0)·sign p) and capsule (u ∝ (x, y', z)); it is piecewise
// constant for the remaining solids.  The capsule's direction also depends on
// its half-length h outside the segment (y' = y -+ h): that term goes to gbase[0].
__device__ __forceinline__ void radius_factor_adjoint(int base, V3 p, const float* prm, V3 s, float gf, V3& gpos,
                                                      V3& gs, float* gbase) {
    V3 u = ref_dir(base, p, prm);
    V3 us = mul(u, s);
    float nus = safe_len(us), nu = safe_len(u);
    if (nus <= 0.f || nu <= 0.f) return;
    float f = nus / nu;
    // explicit scale dependence
    gs = gs + mul(mul(u, u), s) * (gf / (nus * nu));
    // df/du
    V3 dfdu = mul(us, s) * (1.f / (nus * nu)) - u * (f / (nu * nu));
    V3 jt = v3s(0.f);  // (du/dp)ᵀ dfdu for an unnormalised direction (f is scale invariant)
    switch (base) {
        case SPHERE:
            jt = dfdu * (1.f / sqrtf(dot(p, p) + 1e-8f));
            // u = p * rsqrt(|p|^2+eps): the normalisation's Jacobian is radial and
            // does not change f; the tangential part is dfdu / |p|
            break;
        case CUBE:
        case ROUND_CUBE: {
            V3 q = vabs(p) - v3s(1.f);
            if (fmaxf(q.x, fmaxf(q.y, q.z)) > 0.f) {
                // u = k · max(q,0)·sign(p) with k a scalar; du/dp_i = k·[q_i>0]
                float kscale = (base == CUBE) ? 2.f : 1.f / sqrtf(dot(vmax0(q), vmax0(q)) + 1e-8f);
                jt = v3(q.x > 0.f ? dfdu.x * kscale : 0.f, q.y > 0.f ? dfdu.y * kscale : 0.f,
                        q.z > 0.f ? dfdu.z * kscale : 0.f);
            }
            break;
        }
        case CAPSULE: {
            float h = prm[0];
            float y = p.y - clampf(p.y, -h, h);
            float inv = rsqrtf(p.x * p.x + y * y + p.z * p.z + 1e-8f);
            bool inside = (-h <= p.y && p.y <= h);
            jt = v3(dfdu.x * inv, inside ? 0.f : dfdu.y * inv, dfdu.z * inv);
            if (!inside) gbase[0] += gf * dfdu.y * inv * (p.y > 0.f ? -1.f : 1.f);
            break;
        }
        default:
            break;
    }
    gpos = gpos + jt * gf;
}

// ---------------------------------------------------------------------------
// colour fields, strokes_color.h + sh.h (constants from the real-SH normalisation)
// ---------------------------------------------------------------------------
constexpr double INV_SQRT_PI = 0.56418958354775628695;
constexpr float SH_C0 = (float)(0.5 * INV_SQRT_PI);
constexpr float SH_C1 = (float)(0.5 * 1.7320508075688772935 * INV_SQRT_PI);
constexpr float SH_C2 = (float)(0.5 * 3.8729833462074168852 * INV_SQRT_PI);
constexpr float SH_C3 = (float)(0.75 * 2.2360679774997896964 * INV_SQRT_PI);
constexpr float SH_C4 = (float)(0.25 * 2.2360679774997896964 * INV_SQRT_PI);
constexpr float SH_C5 = (float)(0.25 * 3.8729833462074168852 * INV_SQRT_PI);

__host__ __device__ constexpr int num_color_params(int cid) { return cid == 0 ? 3 : cid == 1 ? 12 : cid == 2 ? 12 : 27; }

template <int DEG>
__device__ __forceinline__ void sh_basis_t(V3 dir, float w[9]) {
    float x = dir.x, y = dir.y, z = dir.z;
    w[0] = SH_C0;
    w[1] = -SH_C1 * y;
    w[2] = SH_C1 * z;
    w[3] = -SH_C1 * x;
    if (DEG >= 3) {
        w[4] = SH_C2 * (x * y);
        w[5] = -SH_C2 * (y * z);
        w[6] = SH_C3 * (z * z) - SH_C4;
        w[7] = -SH_C2 * (x * z);
        w[8] = SH_C5 * (x * x) - SH_C5 * (y * y);
    }
}
__device__ __forceinline__ float sigmoidf(float v) { return 1.f / (1.f + expf(-v)); }

template <int DEG>
__device__ __forceinline__ V3 sh_rgb(V3 dir, const float* cp, float w[9]) {
    constexpr int n = DEG == 2 ? 4 : 9;
    sh_basis_t<DEG>(dir, w);
    V3 rgb = v3s(0.f);
#pragma unroll
    for (int i = 0; i < n; i++) rgb = rgb + ld3(cp + 3 * i) * w[i];
    return v3(sigmoidf(rgb.x), sigmoidf(rgb.y), sigmoidf(rgb.z));
}

// colour of a stroke: CID 0 constant, 1 gradient (unit-space position), 2/3 SH (raw view dir)
template <int CID>
__device__ __forceinline__ V3 color_eval_t(V3 pu, V3 dir, const float* cp) {
    if (CID == 0) return ld3(cp);
    if (CID == 1) {
        V3 v0 = pu - ld3(cp), v1 = ld3(cp + 3) - ld3(cp);
        float t = sigmoidf((dot(v0, v1) / dot(v1, v1)) * 5.f - 2.5f);
        return ld3(cp + 6) * (1.f - t) + ld3(cp + 9) * t;
    }
    float w[9];
    return sh_rgb<CID == 2 ? 2 : 3>(dir, cp, w);
}
__device__ __forceinline__ V3 color_eval(int cid, V3 pu, V3 dir, const float* cp) {
    switch (cid) {
        case 0: return color_eval_t<0>(pu, dir, cp);
        case 1: return color_eval_t<1>(pu, dir, cp);
        case 2: return color_eval_t<2>(pu, dir, cp);
        default: return color_eval_t<3>(pu, dir, cp);
    }
}

// adds dL/dcolour-params into gc; returns dL/d(unit pos) (exact mode, gradient_rgb only)
template <int CID>
__device__ __forceinline__ V3 color_grad_t(V3 pu, V3 dir, const float* cp, V3 gcol, float* gc, bool exact) {
    if (CID == 0) { add3(gc, gcol); return v3s(0.f); }
    if (CID == 1) {
        V3 p0 = ld3(cp), p1 = ld3(cp + 3), c0 = ld3(cp + 6), c1 = ld3(cp + 9);
        V3 v0 = pu - p0, v1 = p1 - p0;
        float d0 = dot(v0, v1), d1 = dot(v1, v1);
        float t = sigmoidf((d0 / d1) * 5.f - 2.5f);
        float dt = dot(gcol, c1 - c0);
        float dtr = 5.f * t * (1.f - t) * dt;
        float dd0 = dtr / d1;
        float dd1 = -dtr * d0 / (d1 * d1);
        V3 dv1 = dd0 * v0 + dd1 * 2.f * v1;
        V3 dv0 = dd0 * v1;
        add3(gc + 0, -dv0 - dv1);
        add3(gc + 3, dv1);
        add3(gc + 6, gcol * (1.f - t));
        add3(gc + 9, gcol * t);
        return exact ? dv0 : v3s(0.f);
    }
    constexpr int n = CID == 2 ? 4 : 9;
    float w[9];
    V3 rgb = sh_rgb<CID == 2 ? 2 : 3>(dir, cp, w);
    V3 gr = mul(mul(gcol, rgb), v3s(1.f) - rgb);
#pragma unroll
    for (int i = 0; i < n; i++) add3(gc + 3 * i, gr * w[i]);
    return v3s(0.f);
}

// alpha transfer, strokes_forward.cu:85-89 and common.h:47-51
__device__ __forceinline__ float laplace_cdf(float s) { return s <= 0.f ? 0.5f * expf(s) : 1.f - 0.5f * expf(-s); }

}  // namespace n3ds
