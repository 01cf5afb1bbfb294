// Fused stroke-field renderer: per-sample stroke SDF -> alpha/colour, ordered
// "over" compositing and the volume-render ray march in one pass, with a
// hand-written adjoint that recomputes the SDFs instead of storing per
// (sample, stroke) terms.
//
// Semantics are those of the reference chain for a mixed stroke palette
// (source/models.py:660-711 + source/utils/render.py:5-116): one
// stroke_forward per stroke type placed into columns by stroke index,
// compose_forward over all strokes (index S-1 in front), then
// compute_alpha_weights / volumetric_rendering.
//
// Mapping (one wavefront per ray):
//  * lane l owns samples l, l+64, ... (NS slots, n <= 64*NS); per-sample
//    transmittance / density / colour live in registers and the ray march is
//    a wavefront scan;
//  * strokes are visited in compositing order in chunks of 64: each lane
//    tests one stroke against the ray (unit-space slab test for solids, the
//    boxes of Bezier-hull pieces for curves) and turns the hit interval into
//    a range of sample indices (binary search over the sample depths in LDS);
//  * the hits are compacted with a ballot and their (sample, stroke) pairs
//    are laid out contiguously (wave prefix sum of the range lengths); the
//    expensive per-pair work (SDF, alpha, and in the adjoint the SDF
//    gradient and transform adjoint) runs one pair per lane, so a stroke that
//    covers 5 of 128 samples costs 5 lanes, not a wavefront;
//  * the order-dependent part (over-compositing and its adjoint) then walks
//    the hit strokes in order, each lane updating its own samples from the
//    per-pair values staged in LDS;
//  * adjoint: per-pair gradient slots are reduced per stroke through an LDS
//    transpose (each lane sums one (stroke, slot) segment) into per-block
//    LDS accumulators, flushed once per persistent block.
// Culling is exact for the linear alpha transfer (alpha = 0 outside the
// tested volume); for the Laplace transfer it drops alpha below `cull_eps`.
// Sample depths must be non-decreasing along a ray (as the reference's
// stratified samplers produce); a ray whose depths are not sorted falls back
// to evaluating every hit stroke on all its samples (same result, slower).
#include <mutex>

#include "n3ds_capi.h"
#include "n3ds_stroke_common.h"

namespace n3ds {

constexpr int REC = 128;     // floats per stroke record
constexpr int NPIECE = 16;   // curve pieces bounded for culling
constexpr int FWAVES = 4;    // wavefronts per forward block
constexpr int BWAVES = 8;    // wavefronts per backward block (at most)
constexpr int PBUF = 256;    // (sample, stroke) pairs staged per window (>= samples per ray)

// record layout
enum : int {
    R_BASE = 0, R_FLAGS = 1, R_P = 2, R_NB = 3, R_OFFS = 4, R_OFFA = 5, R_OFFT = 6, R_CURVE = 7,
    R_ROT = 8,      // 9: rotation entries r00..r22
    R_TR = 17,      // 3: translation
    R_SC = 20,      // 3: scale (single scale replicated)
    R_ANG = 23,     // 3: euler angles
    R_BP = 26,      // 2: base params
    R_MINS = 28,    // min scale
    R_E0 = 29,      // 3: unit box half extent
    R_E1 = 32,      // 3: margin multiplier
    R_M = 36,       // 9: culling map diag(1/s) R^T (row-major)
    R_SZ = 45,      // sin / cos of the z angle (Euler adjoint)
    R_CZ = 46,
    R_CP = 8,       // 14: curve params (curves reuse the solid slots)
    R_PTS = 22,     // (NPIECE + 1) * 3: curve points at the piece ends
    R_DEV = 73,     // NPIECE: bound on the distance of a piece from its chord
    R_HULL = 118,   // 6: whole-curve box
    R_RMAX = 124,   // max tube radius
};
static_assert(R_PTS + 3 * (NPIECE + 1) <= R_DEV && R_DEV + NPIECE <= R_HULL && R_RMAX < REC,
              "stroke record layout");

struct RenderArgs {
    const float* origins;
    const float* dirs;
    const float* viewdirs;
    const float* radii;
    const float* tdist;
    const float* table;   // [S, REC]
    const float* cparams; // [S, Q]
    const float* dparams; // [S]
    int64_t R;
    int n, S, Q, Pmax, cid;
    float delta;
    int laplace, isr, cone, exact;
    float coord_div;
    float bg0, bg1, bg2;
    const float* bgs;  // per-ray background colours [R, 3] (nullptr: bg0..bg2 for every ray)
    int opaque;        // opaque_background: the last interval is infinitely wide
    float tail;  // culling margin multiplier
    // hits and per-pair values the forward keeps for the adjoint (hdir == nullptr: none).
    // hdir [R, nch, 3]: per 64-stroke chunk (hit offset or -1 if not kept, pair offset, hits);
    // hbuf [R, cap_h, 4]: stroke, first sample, samples, pair offset within the chunk;
    // pbuf [R, npf, cap_p]: alpha (+ gradient_rgb colour) planes.
    int* hdir;
    int* hbuf;
    float* pbuf;
    int cap_h, cap_p, nch;
    int nch_stride;  // chunks per ray in hdir (the largest scene's in a batch)
    // batch of independent scenes (nullptr: one scene).  Per scene SCENE_REC ints
    // {first ray, rays, first stroke, strokes, first backward block, backward blocks,
    // offset of its partial gradient rows, backward wavefronts per block (0: global
    // accumulators), first culling task, 0}, then the scene of every backward block;
    // rays and strokes of one scene are contiguous in every per-ray / per-stroke array.
    const int* scenes;
    int n_scenes;
    int block_base;  // backward: global block number of this launch's first block
    // hits of the culling pass (nullptr: the forward culls): per (ray, chunk) a count
    // word (hits | pairs << 8) and up to 64 packed hits (pack_hit), index ray * nch_stride + chunk
    const int* ccounts;
    const int* cslots;
    int64_t ntask;    // culling tasks: rays x chunk groups (a batch: each scene's own chunks)
    int cull_group;   // chunks culled by one culling task
    int flush_pairs;  // pending pairs that trigger evaluation
    // split adjoint (render_bwd_kernel MODE 2, then render_bwd_pairs_kernel): per SDF
    // family and MODE 2 wavefront a field-major region of sp_cap adjoint records
    // (stroke, ray, depth, dL/dalpha, dL/d(density), dL/d(colour) x 3), sp_count [NCLASS,
    // sp_nw] the 64-record groups each region holds (then their exclusive prefix and the
    // total, group_scan_kernel), sp_ovf set when a region overflowed
    // (the fused kernel then takes every ray), fbflag [R] the rays left to the fused
    // kernel; nullptr outside the split adjoint
    float* sp_list;
    int* sp_count;
    int* sp_ovf;
    int* fbflag;
    int64_t sp_cap;
    int sp_nw;
};

constexpr int SCENE_REC = 10;

// last scene whose field f is <= v (scenes with nothing in that field share their
// start with the next scene, so the search lands on the one that owns v)
__device__ __forceinline__ int find_scene(const int* sc, int n_scenes, int f, int v) {
    int lo = 0, hi = n_scenes - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (sc[mid * SCENE_REC + f] <= v) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// point the stroke arrays of A at one scene of a batch
__device__ __forceinline__ void scene_strokes(RenderArgs& A, const int* rec) {
    const int sb = rec[2], ns = rec[3];
    A.table += (int64_t)sb * REC;
    A.cparams += (int64_t)sb * A.Q;
    A.dparams += sb;
    A.S = ns;
    A.nch = (ns + 63) / 64;
}

__device__ __forceinline__ int ibits(float f) { return __float_as_int(f); }
__device__ __forceinline__ float fbits(int i) { return __int_as_float(i); }

// ---------------------------------------------------------------------------
// stroke table
// ---------------------------------------------------------------------------
// Bezier control points of the curve piece [a, b]; the piece lies in their convex hull
__device__ void curve_piece_ctrl(int base, const V3 K[4], float a, float b, V3 c[4]) {
    if (base == QUAD_BEZIER) {
        // quadratic P(t) = K0 (1-t)^2 + 2 K1 t(1-t) + K2 t^2
        auto P = [&](float t) { float u = 1.f - t; return (u * u) * K[0] + (2.f * t * u) * K[1] + (t * t) * K[2]; };
        auto dP = [&](float t) { return (2.f * (1.f - t)) * (K[1] - K[0]) + (2.f * t) * (K[2] - K[1]); };
        c[0] = P(a);
        c[3] = P(b);
        c[1] = c[0] + ((b - a) * 0.5f) * dP(a);
        c[2] = c[1];
    } else {
        bool power = base == CATMULL_ROM;
        V3 pa = power ? curve_point<true>(a, K) : curve_point<false>(a, K);
        V3 pb = power ? curve_point<true>(b, K) : curve_point<false>(b, K);
        V3 ta = power ? curve_tangent<true>(a, K) : curve_tangent<false>(a, K);
        V3 tb = power ? curve_tangent<true>(b, K) : curve_tangent<false>(b, K);
        c[0] = pa;
        c[3] = pb;
        c[1] = pa + ((b - a) / 3.f) * ta;
        c[2] = pb - ((b - a) / 3.f) * tb;
    }
}

__device__ float point_segment_dist(V3 p, V3 a, V3 b) {
    const V3 ab = b - a, ap = p - a;
    const float l2 = dot(ab, ab);
    const float h = l2 > 0.f ? fminf(fmaxf(dot(ap, ab) / l2, 0.f), 1.f) : 0.f;
    return safe_len(ap - ab * h);
}

// One stroke per NPIECE consecutive lanes (a curve's pieces one per lane).
__global__ __launch_bounds__(128) void render_prep_kernel(float* __restrict__ table, const int* __restrict__ ids,
                                                          const float* __restrict__ sp, int S, int Pmax,
                                                          int* __restrict__ status) {
    const int64_t gt = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t k = gt / NPIECE;
    const int pl = (int)(gt % NPIECE);
    const bool on = k < S;  // the same for the NPIECE lanes of a stroke
    float* rec = table + (on ? k : 0) * REC;
    if (on)
        for (int i = pl; i < REC; i += NPIECE) rec[i] = 0.f;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    if (!on) return;
    uint32_t id = (uint32_t)ids[k];
    int base = (int)(id >> 4);
    uint32_t flags = id & 15u;
    int P = num_shape_params(id);
    if (base >= NUM_BASE || P > Pmax || ((flags & F_SS) && (flags & F_MS)) || (is_curve(base) && flags)) {
        if (pl == 0) atomicOr(status, 1);
        base = SPHERE; flags = 0; P = 0;
    }
    const float* row = sp + k * Pmax;
    int nb = num_base_params(base);
    if (pl == 0) {
        rec[R_BASE] = fbits(base);
        rec[R_FLAGS] = fbits((int)flags);
        rec[R_P] = fbits(P);
        rec[R_NB] = fbits(nb);
        rec[R_CURVE] = fbits(is_curve(base) ? 1 : 0);
    }
    if (!is_curve(base)) {
        if (pl != 0) return;
        Frame F = load_frame(row, P, flags);
        rec[R_OFFS] = fbits(F.off_scale);
        rec[R_OFFA] = fbits(F.off_angle);
        rec[R_OFFT] = fbits(F.off_trans);
        Rot Rm;
        if (flags & F_R) Rm = F.R;
        else for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++) Rm.r[i][j] = (i == j) ? 1.f : 0.f;
        for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++) rec[R_ROT + 3 * i + j] = Rm.r[i][j];
        rec[R_TR + 0] = F.trans.x; rec[R_TR + 1] = F.trans.y; rec[R_TR + 2] = F.trans.z;
        rec[R_SC + 0] = F.scale.x; rec[R_SC + 1] = F.scale.y; rec[R_SC + 2] = F.scale.z;
        rec[R_ANG + 0] = F.angle.x; rec[R_ANG + 1] = F.angle.y; rec[R_ANG + 2] = F.angle.z;
        rec[R_SZ] = F.sz; rec[R_CZ] = F.cz;
        for (int i = 0; i < nb; i++) rec[R_BP + i] = row[i];
        rec[R_MINS] = fminf(F.scale.x, fminf(F.scale.y, F.scale.z));
        const float isc[3] = {1.f / F.scale.x, 1.f / F.scale.y, 1.f / F.scale.z};
        for (int a = 0; a < 3; a++)
            for (int j = 0; j < 3; j++) rec[R_M + 3 * a + j] = Rm.r[j][a] * isc[a];
        // unit-space box: points with sdf < m satisfy |p_a| < e0_a + e1_a m
        V3 e0 = v3s(1.f), e1 = v3s(1.f);
        float b0 = nb > 0 ? row[0] : 0.f, b1 = nb > 1 ? row[1] : 0.f;
        switch (base) {
            case ROUND_CUBE: e0 = v3s(1.f + b0); break;
            case CAPPED_TORUS: e0 = v3(1.f + b1, 1.f + b1, b1); break;
            case CAPSULE: e0 = v3(1.f, b0 + 1.f, 1.f); break;
            case LINE: { float rr = 1.f + fabsf(b1); e0 = v3(rr, b0 + rr, rr); break; }
            case TRIPRISM: e0 = v3(0.75f / TRI_C, b0, 1.f); e1 = v3(1.5f / TRI_C, 1.f, 2.f); break;
            case OCTAHEDRON:
            case TETRAHEDRON: e1 = v3s(1.7320509f); break;
            default: break;
        }
        rec[R_E0 + 0] = e0.x * 1.0001f + 1e-5f; rec[R_E0 + 1] = e0.y * 1.0001f + 1e-5f; rec[R_E0 + 2] = e0.z * 1.0001f + 1e-5f;
        rec[R_E1 + 0] = e1.x * 1.0001f; rec[R_E1 + 1] = e1.y * 1.0001f; rec[R_E1 + 2] = e1.z * 1.0001f;
    } else {
        if (pl == 0) {
            rec[R_OFFS] = fbits(-1); rec[R_OFFA] = fbits(-1); rec[R_OFFT] = fbits(-1);
            for (int i = 0; i < nb; i++) rec[R_CP + i] = row[i];
        }
        V3 K[4];
        float r1, r2;
        if (base == QUAD_BEZIER) {
            K[0] = ld3(row); K[1] = ld3(row + 3); K[2] = ld3(row + 6); K[3] = K[2];
            r1 = row[9]; r2 = row[10];
        } else if (base == CUBIC_BEZIER) {
            cubic_ctrl(row, K);
            r1 = row[12]; r2 = row[13];
        } else {
            CatmullCoef c = catmull_coef(row);
            for (int j = 0; j < 4; j++) K[j] = c.K[j];
            r1 = row[12]; r2 = row[13];
        }
        // pieces as chord capsules: the piece [j, j+1] / NPIECE lies within dev_j of the chord
        // between its end points (the distance to a segment is convex, so its maximum over the
        // hull of the piece's control points is at a control point); the hull of all control
        // points bounds the whole curve.  Lane j takes piece j.
        const int j = pl;
        V3 c[4];
        curve_piece_ctrl(base, K, (float)j / NPIECE, (float)(j + 1) / NPIECE, c);
        float hull[6] = {1e30f, 1e30f, 1e30f, -1e30f, -1e30f, -1e30f};
        float mag = 0.f;
        for (int q = 0; q < 4; q++) {
            hull[0] = fminf(hull[0], c[q].x); hull[1] = fminf(hull[1], c[q].y); hull[2] = fminf(hull[2], c[q].z);
            hull[3] = fmaxf(hull[3], c[q].x); hull[4] = fmaxf(hull[4], c[q].y); hull[5] = fmaxf(hull[5], c[q].z);
            mag = fmaxf(mag, fmaxf(fabsf(c[q].x), fmaxf(fabsf(c[q].y), fabsf(c[q].z))));
        }
        const float dev = fmaxf(point_segment_dist(c[1], c[0], c[3]), point_segment_dist(c[2], c[0], c[3]));
        rec[R_DEV + j] = dev * 1.001f + 1e-5f * (1.f + mag);
        rec[R_PTS + 3 * j + 0] = c[0].x; rec[R_PTS + 3 * j + 1] = c[0].y; rec[R_PTS + 3 * j + 2] = c[0].z;
        if (j == NPIECE - 1) {
            rec[R_PTS + 3 * NPIECE + 0] = c[3].x;
            rec[R_PTS + 3 * NPIECE + 1] = c[3].y;
            rec[R_PTS + 3 * NPIECE + 2] = c[3].z;
        }
        // min / max over the stroke's NPIECE lanes (aligned groups, so xor stays inside)
#pragma unroll
        for (int off = 1; off < NPIECE; off <<= 1) {
#pragma unroll
            for (int a = 0; a < 3; a++) {
                hull[a] = fminf(hull[a], __shfl_xor(hull[a], off, 64));
                hull[3 + a] = fmaxf(hull[3 + a], __shfl_xor(hull[3 + a], off, 64));
            }
        }
        if (pl == 0) {
            for (int a = 0; a < 3; a++) {
                const float pad = 1e-5f * (1.f + fmaxf(fabsf(hull[a]), fabsf(hull[3 + a])));
                hull[a] -= pad;
                hull[3 + a] += pad;
            }
            for (int a = 0; a < 6; a++) rec[R_HULL + a] = hull[a];
            rec[R_RMAX] = fmaxf(fmaxf(r1, r2), 0.f) * 1.0001f + 1e-6f;
        }
    }
}

// ---------------------------------------------------------------------------
// culling: parametric interval of the ray segment inside the stroke's bound
// ---------------------------------------------------------------------------
__device__ __forceinline__ void slab(float o, float d, float lo, float hi, float& t0, float& t1) {
    float inv = 1.f / d;
    float a = (lo - o) * inv, b = (hi - o) * inv;
    t0 = fmaxf(t0, fminf(a, b));
    t1 = fminf(t1, fmaxf(a, b));
}

// slab test with a precomputed inverse direction component
__device__ __forceinline__ void slab_inv(float o, float inv, float lo, float hi, float& t0, float& t1) {
    float a = (lo - o) * inv, b = (hi - o) * inv;
    t0 = fmaxf(t0, fminf(a, b));
    t1 = fminf(t1, fmaxf(a, b));
}

__device__ __forceinline__ void pad_interval(float t0, float t1, float& ta, float& tb) {
    float pad = 1e-4f * (t1 - t0) + 1e-5f * (fabsf(t0) + fabsf(t1)) + 1e-6f;
    ta = t0 - pad;
    tb = t1 + pad;
}

// Order-preserving map of floats to ints (for LDS integer min / max) and back.
__device__ __forceinline__ int ord_int(float f) {
    const int i = __float_as_int(f);
    return i >= 0 ? i : i ^ 0x7fffffff;
}
__device__ __forceinline__ float ord_float(int i) { return __int_as_float(i >= 0 ? i : i ^ 0x7fffffff); }

// Ray interval [t0, t1] within distance e of the segment [a, b] (a capsule, so the
// interval is convex): hull of the intervals in the two end spheres and in the
// cylinder part.  idd = 1 / |d|^2.  Hardware reciprocals and square roots (1 ulp);
// e is padded relatively so rounding cannot drop a boundary point.
__device__ __forceinline__ bool capsule_interval(V3 o, V3 d, float idd, V3 a, V3 b, float e, float& t0, float& t1) {
    e = e * 1.0002f + 1e-6f;
    const float e2 = e * e;
    const V3 ba = b - a, oa = o - a, ob1 = o - b;
    const float l2 = dot(ba, ba), dd = dot(d, d);
    const float db = dot(d, ba), ob = dot(oa, ba);
    float lo = 1e30f, hi = -1e30f;
    // end spheres
#pragma unroll
    for (int end = 0; end < 2; end++) {
        const V3 q0 = end ? ob1 : oa;
        const float tc = -dot(q0, d) * idd;
        const V3 q = q0 + d * tc;
        const float h = e2 - dot(q, q);
        if (h >= 0.f) {
            const float w = __builtin_amdgcn_sqrtf(h * idd);
            lo = fminf(lo, tc - w);
            hi = fmaxf(hi, tc + w);
        }
    }
    // cylinder around the chord, between the planes through a and b
    if (l2 > 0.f) {
        const float il2 = __builtin_amdgcn_rcpf(l2);
        const V3 dp = d - ba * (db * il2), op = oa - ba * (ob * il2);  // components normal to the axis
        const float A2 = dot(dp, dp), B2 = dot(op, dp), C2 = dot(op, op) - e2;
        float c0 = -1e30f, c1 = 1e30f;
        bool any = true;
        if (A2 > 1e-12f * dd) {
            const float disc = B2 * B2 - A2 * C2;
            if (disc >= 0.f) {
                const float sq = __builtin_amdgcn_sqrtf(disc), ia = __builtin_amdgcn_rcpf(A2);
                c0 = (-B2 - sq) * ia;
                c1 = (-B2 + sq) * ia;
            } else {
                any = false;
            }
        } else {
            any = C2 <= 0.f;
        }
        if (any) {
            if (fabsf(db) > 1e-12f * sqrtf(dd * l2)) {
                const float idb = __builtin_amdgcn_rcpf(db);
                const float s0 = -ob * idb, s1 = (l2 - ob) * idb;
                c0 = fmaxf(c0, fminf(s0, s1));
                c1 = fminf(c1, fmaxf(s0, s1));
            } else if (ob < 0.f || ob > l2) {
                any = false;
            }
            if (any && c0 <= c1) {
                lo = fminf(lo, c0);
                hi = fmaxf(hi, c1);
            }
        }
    }
    t0 = lo;
    t1 = hi;
    return lo <= hi;
}

// Ray interval [ta, tb] inside the stroke's bound.  Solids: final.  Curves
// (*curve = true): the whole-curve box only; the piece boxes refine it
// (refine_curve_hits), one lane per piece.
// di: componentwise 1 / d (curves are tested in world space, so it is per ray)
__device__ __forceinline__ bool cull_stroke(const float* __restrict__ rec, V3 o, V3 d, V3 di, float m, int isr,
                                            float tlo, float thi, float& ta, float& tb, bool& curve) {
    float t0 = tlo, t1 = thi;
    curve = ibits(rec[R_CURVE]) != 0;
    if (curve) {
        float e = rec[R_RMAX] + m;
        slab_inv(o.x, di.x, rec[R_HULL + 0] - e, rec[R_HULL + 3] + e, t0, t1);
        slab_inv(o.y, di.y, rec[R_HULL + 1] - e, rec[R_HULL + 4] + e, t0, t1);
        slab_inv(o.z, di.z, rec[R_HULL + 2] - e, rec[R_HULL + 5] + e, t0, t1);
        ta = t0;
        tb = t1;
        return t0 <= t1;
    } else {
        // ray in the stroke's unit space: A + t B = M (o + t d - tr), M = diag(1/s) R^T; the
        // reciprocals are the hardware approximation (1 ulp), far inside the padding below
        const float* M = rec + R_M;
        V3 q = o - ld3(rec + R_TR);
        V3 A = v3(M[0] * q.x + M[1] * q.y + M[2] * q.z, M[3] * q.x + M[4] * q.y + M[5] * q.z,
                  M[6] * q.x + M[7] * q.y + M[8] * q.z);
        V3 B = v3(M[0] * d.x + M[1] * d.y + M[2] * d.z, M[3] * d.x + M[4] * d.y + M[5] * d.z,
                  M[6] * d.x + M[7] * d.y + M[8] * d.z);
        float mu = isr ? m / rec[R_MINS] : m;
        float ex = rec[R_E0 + 0] + rec[R_E1 + 0] * mu;
        float ey = rec[R_E0 + 1] + rec[R_E1 + 1] * mu;
        float ez = rec[R_E0 + 2] + rec[R_E1 + 2] * mu;
        slab_inv(A.x, __builtin_amdgcn_rcpf(B.x), -ex, ex, t0, t1);
        slab_inv(A.y, __builtin_amdgcn_rcpf(B.y), -ey, ey, t0, t1);
        slab_inv(A.z, __builtin_amdgcn_rcpf(B.z), -ez, ez, t0, t1);
        const int base = ibits(rec[R_BASE]);
        if (base == OCTAHEDRON || base == TETRAHEDRON) {
            // face planes: sdf < mu implies |x|+|y|+|z| < ex (octahedron) and
            // max(|x+y| - z, |x-y| + z) < ex (tetrahedron), ex = the box's padded extent;
            // the sums get a rounding pad of their own
            const float c = ex + 4e-7f * (fabsf(A.x) + fabsf(A.y) + fabsf(A.z));
            const float pa = A.x + A.y - A.z, pb = B.x + B.y - B.z;  // x + y - z
            const float qa = A.x - A.y + A.z, qb = B.x - B.y + B.z;  // x - y + z
            const float ra = A.y + A.z - A.x, rb = B.y + B.z - B.x;  // -x + y + z
            const float sa = A.x + A.y + A.z, sb = B.x + B.y + B.z;  // x + y + z
            if (base == OCTAHEDRON) {
                slab_inv(pa, __builtin_amdgcn_rcpf(pb), -c, c, t0, t1);
                slab_inv(qa, __builtin_amdgcn_rcpf(qb), -c, c, t0, t1);
                slab_inv(ra, __builtin_amdgcn_rcpf(rb), -c, c, t0, t1);
                slab_inv(sa, __builtin_amdgcn_rcpf(sb), -c, c, t0, t1);
            } else {
                slab_inv(pa, __builtin_amdgcn_rcpf(pb), -1e30f, c, t0, t1);
                slab_inv(qa, __builtin_amdgcn_rcpf(qb), -1e30f, c, t0, t1);
                slab_inv(ra, __builtin_amdgcn_rcpf(rb), -1e30f, c, t0, t1);
                slab_inv(-sa, __builtin_amdgcn_rcpf(-sb), -1e30f, c, t0, t1);  // -x - y - z
            }
        }
    }
    if (!(t0 <= t1)) return false;
    pad_interval(t0, t1, ta, tb);
    return true;
}

// ---------------------------------------------------------------------------
// per (sample, stroke) evaluation
// ---------------------------------------------------------------------------
struct Pair {
    V3 p;
    float sdf, alpha, sc, fac;
};

__device__ __forceinline__ Pair eval_pair(const float* __restrict__ rec, int base, bool curve, V3 x, float r,
                                          const RenderArgs& A) {
    Pair q;
    const float* prm = curve ? rec + R_CP : rec + R_BP;
    if (curve) {
        q.p = x;
    } else {
        // unit-space point diag(1/s) R^T (x - t) through the stored map (no divisions)
        const float* m = rec + R_M;
        V3 d = x - ld3(rec + R_TR);
        q.p = v3(m[0] * d.x + m[1] * d.y + m[2] * d.z, m[3] * d.x + m[4] * d.y + m[5] * d.z,
                 m[6] * d.x + m[7] * d.y + m[8] * d.z);
    }
    q.sdf = sdf_eval(base, q.p, prm);
    q.fac = 1.f;
    if (A.isr && !curve) {
        uint32_t flags = (uint32_t)ibits(rec[R_FLAGS]);
        if (flags & F_SS) { q.fac = rec[R_SC]; r = r / q.fac; }
        else if (flags & F_MS) { q.fac = radius_factor(ref_dir(base, q.p, prm), ld3(rec + R_SC)); r = r / q.fac; }
    }
    q.sc = (A.laplace ? 2.f : 0.5f) / (A.delta * r);
    q.alpha = alpha_from_sdf(q.sdf, q.sc, A.delta, A.laplace != 0);
    return q;
}

// dL/dSDF of the exact adjoint as a function of the SDF value (alpha recomputed from it)
struct ExactTransfer {
    float ga, sc, delta;
    bool laplace;
    __device__ __forceinline__ float operator()(float sdf) const {
        const float a = alpha_from_sdf(sdf, sc, delta, laplace);
        return ga * dalpha_dsdf(sdf, a, sc, delta, laplace, true);
    }
};

template <int CID>
__device__ __forceinline__ V3 stroke_color_uniform(const float* __restrict__ cp, V3 vdir) {
    if (CID == 0) return ld3(cp);
    if (CID >= 2) return color_eval_t<CID>(v3s(0.f), vdir, cp);
    return v3s(0.f);
}

// ---------------------------------------------------------------------------
// per-ray constants and the per-wavefront pair machinery
// ---------------------------------------------------------------------------
struct RayConst {
    V3 o, d, vd, oc, dc;  // oc, dc: origin / direction in the (warped) stroke coordinates
    V3 di;                // 1 / dc componentwise
    float radii, tlo, thi, m;
    float idd;            // 1 / |dc|^2
};

__device__ __forceinline__ RayConst load_ray(const RenderArgs& A, int64_t ray) {
    RayConst c;
    c.o = ld3(A.origins + 3 * ray);
    c.d = ld3(A.dirs + 3 * ray);
    c.vd = ld3(A.viewdirs + 3 * ray);
    c.radii = A.radii[ray];
    const float* td = A.tdist + ray * (A.n + 1);
    c.tlo = 0.5f * (td[0] + td[1]);
    c.thi = 0.5f * (td[A.n - 1] + td[A.n]);
    const float rmax = A.cone ? c.radii * fmaxf(fabsf(c.tlo), fabsf(c.thi)) : c.radii;
    c.m = A.delta > 0.f ? A.delta * rmax * A.tail : 1e-6f;
    c.oc = c.o;
    c.dc = c.d;
    if (A.coord_div != 1.f) {
        c.oc = v3(c.o.x / A.coord_div, c.o.y / A.coord_div, c.o.z / A.coord_div);
        c.dc = v3(c.d.x / A.coord_div, c.d.y / A.coord_div, c.d.z / A.coord_div);
    }
    c.di = v3(1.f / c.dc.x, 1.f / c.dc.y, 1.f / c.dc.z);
    c.idd = 1.f / dot(c.dc, c.dc);
    return c;
}

// sample position (stroke coordinates) and radius at depth t, render.py:5-26
__device__ __forceinline__ void sample_point(const RenderArgs& A, const RayConst& c, float t, V3& x, float& r) {
    V3 xw = c.d * t + c.o;
    if (A.coord_div != 1.f) xw = v3(xw.x / A.coord_div, xw.y / A.coord_div, xw.z / A.coord_div);
    x = xw;
    r = A.cone ? c.radii * t : c.radii;
}

// LDS ordering between the lanes of one wavefront (no block barrier)
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ int wave_incl_scan(int v, int lane) {
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        int u = __shfl_up(v, off, 64);
        if (lane >= off) v += u;
    }
    return v;
}

// Hits are gathered over consecutive 64-stroke chunks (in compositing order)
// until HMAX - 64 hits or FLUSH_PAIRS pairs are pending, so that the
// one-pair-per-lane batches are full even though a chunk typically covers
// only a few (sample, stroke) pairs of a ray.
constexpr int HMAX = 128;         // pending hits per wavefront
constexpr int FLUSH_PAIRS = 128;  // pending pairs that trigger evaluation (default; tunable)

// per-wavefront LDS region
struct WaveMem {
    float* ts;   // [64*NS] sample depth, +inf past the last sample
    int* ck;     // [HMAX] stroke index of the c-th pending hit (compositing order)
    int* ci0;    // [HMAX] first sample of its interval
    int* coff;   // [HMAX] offset of its first pair among the pending pairs
    int* ccnt;   // [HMAX] its number of pairs
    int* ccls;   // [HMAX] its stroke class (SDF family, see stroke_class)
    int* xo;     // [64] class lists: first pair of the r-th hit of one class ...
    int* xh;     // [64] ... and that hit
    float* pb;   // [npf][PBUF] per-pair values of the current window
    float* scr;  // [64][ng+1] per-lane gradient slots (adjoint)
    int* ul;     // [PBUF] window pairs with a nonzero adjoint (adjoint)
};

__host__ __device__ constexpr int wave_mem_floats(int ns, int npf, int ng) {
    return 64 * ns + 5 * HMAX + 128 + npf * PBUF + (ng > 0 ? 64 * (ng + 1) + PBUF : 0);
}

__device__ __forceinline__ WaveMem wave_mem(float* base, int ns, int npf, int ng = 0) {
    WaveMem w;
    w.ts = base;
    int* ib = (int*)(base + 64 * ns);
    w.ck = ib;
    w.ci0 = ib + HMAX;
    w.coff = ib + 2 * HMAX;
    w.ccnt = ib + 3 * HMAX;
    w.ccls = ib + 4 * HMAX;
    w.xo = ib + 5 * HMAX;
    w.xh = ib + 5 * HMAX + 64;
    w.pb = base + 64 * ns + 5 * HMAX + 128;
    w.scr = w.pb + npf * PBUF;
    w.ul = (int*)(w.scr + (ng > 0 ? 64 * (ng + 1) : 0));
    return w;
}

// number of samples with depth < v (LE: <= v); ts is sorted with 64*NS entries
template <int NS, bool LE>
__device__ __forceinline__ int count_below(const float* ts, float v) {
    int pos = 0;
#pragma unroll
    for (int step = 256; step > 0; step >>= 1) {
        if (pos + step <= 64 * NS) {
            const float t = ts[pos + step - 1];
            if (LE ? (t <= v) : (t < v)) pos += step;
        }
    }
    return pos;
}

// hit c in [c0, c1) whose pair range contains window-relative pair p
__device__ __forceinline__ int find_hit(const WaveMem& w, int c0, int c1, int wbase, int p) {
    int c = c0;
#pragma unroll
    for (int step = 32; step > 0; step >>= 1) {
        const int m = c + step;
        if (m < c1 && w.coff[m] - wbase <= p) c = m;
    }
    return c;
}

// Pairs are evaluated class by class so that a 64-pair group runs one SDF
// family: 0 cubic Bezier / Catmull-Rom, 1 quadratic Bezier, 2 solids.
constexpr int NCLASS = 3;

__device__ __forceinline__ int stroke_class(const float* rec, bool curve) {
    return curve ? (ibits(rec[R_BASE]) == QUAD_BEZIER ? 1 : 0) : 2;
}

// Orders the window's hits by class (stable): w.xh[r] = r-th hit of the order,
// w.xo[r] = its first pair in the class-ordered pair sequence.  Lane h holds hit
// hc = c0 + h of class hcls (-1 past the window).  Returns the number of hits.
__device__ __forceinline__ int class_sort_hits(const WaveMem& w, int hc, int hcls, int lane) {
    const int cnt = hcls >= 0 ? w.ccnt[hc] : 0;
    int before_hits = 0, before_pairs = 0, rank = 0, xo = 0;
#pragma unroll
    for (int x = 0; x < NCLASS; x++) {
        const bool in = hcls == x;
        const unsigned long long m = __ballot(in);
        const int incl = wave_incl_scan(in ? cnt : 0, lane);
        if (in) {
            rank = before_hits + __popcll(m & ((1ull << lane) - 1ull));
            xo = before_pairs + incl - cnt;
        }
        before_hits += __popcll(m);
        before_pairs += lane_i(incl, 63);
    }
    if (hcls >= 0) {
        w.xo[rank] = xo;
        w.xh[rank] = hc;
    }
    wave_sync();
    return before_hits;
}

// r-th hit of the current class list holding class pair pp
__device__ __forceinline__ int class_find(const WaveMem& w, int nx, int pp) {
    int r = 0;
#pragma unroll
    for (int step = 32; step > 0; step >>= 1)
        if (r + step < nx && w.xo[r + step] <= pp) r += step;
    return r;
}

// Writes the ray's sample depths to LDS and returns whether they are sorted.
template <int NS>
__device__ __forceinline__ bool stage_depths(const RenderArgs& A, int64_t ray, int lane, const WaveMem& w,
                                             float (&t)[NS], bool (&valid)[NS]) {
    const float* td = A.tdist + ray * (A.n + 1);
#pragma unroll
    for (int s = 0; s < NS; s++) {
        const int i = lane + 64 * s;
        valid[s] = i < A.n;
        const int ii = valid[s] ? i : 0;
        t[s] = (td[ii] + td[ii + 1]) / 2.f;
        w.ts[i] = valid[s] ? t[s] : INFINITY;
    }
    wave_sync();
    bool bad = false;
#pragma unroll
    for (int s = 0; s < NS; s++) {
        const int i = lane + 64 * s;
        if (i + 1 < A.n && w.ts[i] > w.ts[i + 1]) bad = true;
    }
    return __ballot(bad) == 0ull;
}

// Unsorted depths: the segment the strokes are culled against spans the smallest and
// largest sample depth (not the first and last), and the cone margin follows it.
template <int NS>
__device__ __forceinline__ void unsorted_range(const RenderArgs& A, RayConst& rc, const float (&ts)[NS],
                                               const bool (&valid)[NS]) {
    float lo = INFINITY, hi = -INFINITY;
#pragma unroll
    for (int s = 0; s < NS; s++)
        if (valid[s]) {
            lo = fminf(lo, ts[s]);
            hi = fmaxf(hi, ts[s]);
        }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        lo = fminf(lo, __shfl_xor(lo, off, 64));
        hi = fmaxf(hi, __shfl_xor(hi, off, 64));
    }
    rc.tlo = lo;
    rc.thi = hi;
    const float rmax = A.cone ? rc.radii * fmaxf(fabsf(lo), fabsf(hi)) : rc.radii;
    rc.m = A.delta > 0.f ? A.delta * rmax * A.tail : 1e-6f;
}

// Culls the 64 strokes of the chunk starting at base0 (in compositing order),
// one per lane: stroke k, whether it is hit, its samples [i0, i0 + cnt), its
// class and, for a curve, the mask of its pieces whose chord boxes pass (the
// samples then cover the union of those boxes, to be narrowed by capsule_range).
template <int NS>
__device__ __forceinline__ bool cull_chunk(const RenderArgs& A, const RayConst& rc, const WaveMem& w, bool sorted,
                                           int base0, int lane, int& k, int& i0, int& cnt, int& cls, int& pmask) {
    k = A.S - 1 - (base0 + lane);
    float ta = 0.f, tb = -1.f;
    bool curve = false;
    bool hit = k >= 0 && cull_stroke(A.table + (int64_t)k * REC, rc.oc, rc.dc, rc.di, rc.m, A.isr, rc.tlo, rc.thi,
                                     ta, tb, curve);
    // curves whose whole-curve box is hit: their NPIECE pieces are tested against the
    // piece's chord box with one lane per piece (64 / NPIECE curves at a time).  The
    // hit carries the union of the passing boxes as a provisional interval and their
    // mask; refine_pending narrows it to the chord capsules when the pending hits are
    // evaluated (many chunks at once).  w.pb is free outside the evaluation windows.
    static_assert(NPIECE == 16, "one 16-lane group per curve");
    const unsigned long long cm = __ballot(hit && curve);
    pmask = 0;
    if (cm) {
        int* src = (int*)w.pb;      // [64] lane of the c-th curve
        float* res = w.pb + 64;     // [64][2] box-union interval
        int* msk = (int*)w.pb + 192;  // [64] passing pieces
        const int crank = __popcll(cm & ((1ull << lane) - 1ull));
        if (hit && curve) src[crank] = lane;
        wave_sync();
        const int ncv = __popcll(cm);
        const int kb = A.S - 1 - base0;  // stroke of lane 0 of this chunk
        for (int q0 = 0; q0 < ncv; q0 += 64 / NPIECE) {
            const int cs = q0 + lane / NPIECE;
            const int piece = lane % NPIECE;
            float u0 = 1e30f, u1 = -1e30f;
            bool bh = false;
            if (cs < ncv) {
                const float* rec = A.table + (int64_t)(kb - src[cs]) * REC;
                const float e = rec[R_RMAX] + rc.m + rec[R_DEV + piece];
                const V3 a = ld3(rec + R_PTS + 3 * piece), b = ld3(rec + R_PTS + 3 * piece + 3);
                float p0 = rc.tlo, p1 = rc.thi;
                slab_inv(rc.oc.x, rc.di.x, fminf(a.x, b.x) - e, fmaxf(a.x, b.x) + e, p0, p1);
                slab_inv(rc.oc.y, rc.di.y, fminf(a.y, b.y) - e, fmaxf(a.y, b.y) + e, p0, p1);
                slab_inv(rc.oc.z, rc.di.z, fminf(a.z, b.z) - e, fmaxf(a.z, b.z) + e, p0, p1);
                bh = p0 <= p1;
                if (bh) { u0 = p0; u1 = p1; }
            }
            const unsigned long long bm = __ballot(bh);
#pragma unroll
            for (int off = NPIECE / 2; off > 0; off >>= 1) {
                u0 = fminf(u0, __shfl_xor(u0, off, 64));
                u1 = fmaxf(u1, __shfl_xor(u1, off, 64));
            }
            if (cs < ncv && piece == 0) {
                res[2 * cs] = u0;
                res[2 * cs + 1] = u1;
                msk[cs] = (int)((bm >> (lane & ~(NPIECE - 1))) & 0xffffull);
            }
        }
        wave_sync();
        if (hit && curve) {
            const float u0 = res[2 * crank], u1 = res[2 * crank + 1];
            pmask = msk[crank];
            hit = u0 <= u1;
            if (hit) pad_interval(u0, u1, ta, tb);
        }
        wave_sync();
    }
    i0 = 0;
    cnt = 0;
    cls = 2;
    if (hit) {
        if (sorted) {
            i0 = count_below<NS, false>(w.ts, ta);
            cnt = count_below<NS, true>(w.ts, tb) - i0;
        } else {
            cnt = A.n;
        }
        hit = cnt > 0;
        cls = stroke_class(A.table + (int64_t)k * REC, curve);
    }
    return hit;
}

// Samples [i0, i0 + cnt) of curve stroke k within the chord capsules of its pieces in
// pmask (each piece tested in turn by the calling lane); cnt = 0 if none is hit.
template <int NS>
__device__ __forceinline__ void capsule_range(const RenderArgs& A, const RayConst& rc, const float* ts, bool sorted,
                                              int k, unsigned pmask, int& i0, int& cnt) {
    const float* rec = A.table + (int64_t)k * REC;
    const float e0 = rec[R_RMAX] + rc.m;
    float u0 = 1e30f, u1 = -1e30f;
    for (unsigned m = pmask; m; m &= m - 1) {
        const int pc = __builtin_ctz(m);
        float p0, p1;
        if (capsule_interval(rc.oc, rc.dc, rc.idd, ld3(rec + R_PTS + 3 * pc), ld3(rec + R_PTS + 3 * pc + 3),
                             e0 + rec[R_DEV + pc], p0, p1)) {
            u0 = fminf(u0, fmaxf(p0, rc.tlo));
            u1 = fmaxf(u1, fminf(p1, rc.thi));
        }
    }
    cnt = 0;
    if (u0 <= u1) {
        float ta, tb;
        pad_interval(u0, u1, ta, tb);
        if (sorted) {
            i0 = count_below<NS, false>(ts, ta);
            cnt = count_below<NS, true>(ts, tb) - i0;
        } else {
            i0 = 0;
            cnt = A.n;
        }
    }
}

// Culls a chunk (cull_chunk) and appends the compacted hits to the pending list
// after its nh hits / np pairs; curve hits stay provisional (class | 4 | pieces
// << 16) until refine_pending.  Returns the number of new hits; tot receives
// their number of pairs.
template <int NS>
__device__ __forceinline__ int append_hits(const RenderArgs& A, const RayConst& rc, const WaveMem& w, bool sorted,
                                           int base0, int lane, int nh, int np, int& tot) {
    int k, i0, cnt, cls, pmask;
    const bool hit = cull_chunk<NS>(A, rc, w, sorted, base0, lane, k, i0, cnt, cls, pmask);
    const bool curve = cls < 2;
    const unsigned long long mask = __ballot(hit);
    const int incl = wave_incl_scan(hit ? cnt : 0, lane);
    if (hit) {
        const int rank = nh + __popcll(mask & ((1ull << lane) - 1ull));
        w.ck[rank] = k;
        w.ci0[rank] = i0;
        w.coff[rank] = np + incl - cnt;
        w.ccnt[rank] = cnt;
        // class | provisional (curve still to refine) << 2 | passing pieces << 16
        w.ccls[rank] = cls | (curve ? 4 | (pmask << 16) : 0);
    }
    tot = lane_i(incl, 63);
    wave_sync();
    return __popcll(mask);
}

// Largest window [c0, c1) of pending hits (c1 <= nh, c1 - c0 <= 64) whose pairs fit the staging buffer.
__device__ __forceinline__ int window_end(const WaveMem& w, int c0, int nh, int wbase, int lane) {
    const int c = c0 + lane;
    const bool fits = c < nh && w.coff[c] + w.ccnt[c] - wbase <= PBUF;
    return c0 + __popcll(__ballot(fits));
}

// The window's hits held one per lane (lane h: hit c0 + h) with their stroke's
// density and uniform colour loaded in parallel, so the ordered loops read them
// with v_readlane instead of a dependent LDS + scalar-memory round trip per hit.
template <int CID>
struct WindowHits {
    int k, ci, cn, po;
    float dk;
    V3 cu;
};

template <int CID>
__device__ __forceinline__ WindowHits<CID> window_hits(const RenderArgs& A, const WaveMem& w, int c0, int c1,
                                                      int wbase, int lane, V3 vd) {
    WindowHits<CID> h;
    h.k = h.ci = h.cn = h.po = 0;
    h.dk = 0.f;
    h.cu = v3s(0.f);
    const int c = c0 + lane;
    if (c < c1) {
        h.k = w.ck[c];
        h.ci = w.ci0[c];
        h.cn = w.ccnt[c];
        h.po = w.coff[c] - wbase;
        h.dk = A.dparams[h.k];
        h.cu = stroke_color_uniform<CID>(A.cparams + (int64_t)h.k * A.Q, vd);
    }
    return h;
}

// Number of pending hits (in compositing order) whose stroke lies in a chunk
// before chunk cc; top = index of the first stroke in compositing order.
__device__ __forceinline__ int first_hit_of_chunk(const WaveMem& w, int nh, int top, int cc) {
    int lo = 0;
#pragma unroll
    for (int step = HMAX; step > 0; step >>= 1)
        if (lo + step <= nh && ((top - w.ck[lo + step - 1]) >> 6) < cc) lo += step;
    return lo;
}

// Narrows the provisional curve hits among the nh pending hits to their chord
// capsules (each hit lane tests its passing pieces one after another), recomputes
// their samples, drops hits left without samples and recomputes the pair offsets.
// Returns the new number of pending hits; np receives the pending pairs.
template <int NS>
__device__ __forceinline__ int refine_pending(const RenderArgs& A, const RayConst& rc, const WaveMem& w, bool sorted,
                                              int nh, int lane, int& np) {
    int out = 0, pairs = 0;
    for (int h0 = 0; h0 < nh; h0 += 64) {
        const int c = h0 + lane;
        int k = 0, i0 = 0, cnt = 0, cls = 0;
        if (c < nh) {
            k = w.ck[c];
            i0 = w.ci0[c];
            cnt = w.ccnt[c];
            cls = w.ccls[c];
        }
        if (c < nh && (cls & 4)) {
            capsule_range<NS>(A, rc, w.ts, sorted, k, (unsigned)cls >> 16, i0, cnt);
            cls &= 3;
        }
        const bool keep = c < nh && cnt > 0;
        const unsigned long long km = __ballot(keep);
        const int incl = wave_incl_scan(keep ? cnt : 0, lane);
        wave_sync();  // every lane has read its entry before any is overwritten (new index <= old)
        if (keep) {
            const int r = out + __popcll(km & ((1ull << lane) - 1ull));
            w.ck[r] = k;
            w.ci0[r] = i0;
            w.ccnt[r] = cnt;
            w.ccls[r] = cls;
            w.coff[r] = pairs + incl - cnt;
        }
        out += __popcll(km);
        pairs += lane_i(incl, 63);
        wave_sync();
    }
    np = pairs;
    return out;
}

// ---------------------------------------------------------------------------
// culling pass: one wavefront per (ray, chunk of 64 strokes)
// ---------------------------------------------------------------------------

// cull_chunk with final curve intervals: every piece of a hull-hit curve is tested
// against its chord capsule directly, one lane per piece (no box prefilter: the 16
// lanes of a curve run the capsule math anyway), and the union of the hit
// capsules is the curve's interval.
template <int NS>
__device__ __forceinline__ bool cull_chunk_final(const RenderArgs& A, const RayConst& rc, const WaveMem& w,
                                                 bool sorted, int base0, int lane, int& k, int& i0, int& cnt,
                                                 int& cls) {
    k = A.S - 1 - (base0 + lane);
    float ta = 0.f, tb = -1.f;
    bool curve = false;
    bool hit = k >= 0 && cull_stroke(A.table + (int64_t)k * REC, rc.oc, rc.dc, rc.di, rc.m, A.isr, rc.tlo, rc.thi,
                                     ta, tb, curve);
    const unsigned long long cm = __ballot(hit && curve);
    if (cm) {
        int* src = (int*)w.pb;   // [64] lane of the c-th curve
        float* res = w.pb + 64;  // [64][2] curve interval
        const int crank = __popcll(cm & ((1ull << lane) - 1ull));
        if (hit && curve) src[crank] = lane;
        wave_sync();
        const int ncv = __popcll(cm);
        const int kb = A.S - 1 - base0;
        for (int q0 = 0; q0 < ncv; q0 += 64 / NPIECE) {
            const int cs = q0 + lane / NPIECE;
            const int piece = lane % NPIECE;
            float u0 = 1e30f, u1 = -1e30f;
            if (cs < ncv) {
                const float* rec = A.table + (int64_t)(kb - src[cs]) * REC;
                float p0, p1;
                if (capsule_interval(rc.oc, rc.dc, rc.idd, ld3(rec + R_PTS + 3 * piece),
                                     ld3(rec + R_PTS + 3 * piece + 3), rec[R_RMAX] + rc.m + rec[R_DEV + piece], p0,
                                     p1)) {
                    p0 = fmaxf(p0, rc.tlo);
                    p1 = fminf(p1, rc.thi);
                    if (p0 <= p1) {
                        u0 = p0;
                        u1 = p1;
                    }
                }
            }
#pragma unroll
            for (int off = NPIECE / 2; off > 0; off >>= 1) {
                u0 = fminf(u0, __shfl_xor(u0, off, 64));
                u1 = fmaxf(u1, __shfl_xor(u1, off, 64));
            }
            if (cs < ncv && piece == 0) {
                res[2 * cs] = u0;
                res[2 * cs + 1] = u1;
            }
        }
        wave_sync();
        if (hit && curve) {
            const float u0 = res[2 * crank], u1 = res[2 * crank + 1];
            hit = u0 <= u1;
            if (hit) pad_interval(u0, u1, ta, tb);
        }
        wave_sync();
    }
    i0 = 0;
    cnt = 0;
    cls = 2;
    if (hit) {
        if (sorted) {
            i0 = count_below<NS, false>(w.ts, ta);
            cnt = count_below<NS, true>(w.ts, tb) - i0;
        } else {
            cnt = A.n;
        }
        hit = cnt > 0;
        cls = stroke_class(A.table + (int64_t)k * REC, curve);
    }
    return hit;
}
constexpr int CWAVES = 4;  // wavefronts per culling block
constexpr int CBATCH = 4;  // chunks whose curve pieces are refined together

// per wave: depths [64 NS]; per (batch chunk, lane) interval lo / hi (order-preserving
// ints) and state (hit | curve << 1 | passing pieces << 16); curve -> lane map [64];
// capsule candidates [128]
__host__ __device__ constexpr int cull_wave_floats(int ns) { return 64 * ns + CBATCH * 64 * 3 + 64 + 128; }

// lane in chunk | first sample << 6 | samples << 15 | class << 24 (samples <= 256)
__device__ __forceinline__ int pack_hit(int l, int i0, int cnt, int cls) {
    return l | (i0 << 6) | (cnt << 15) | (cls << 24);
}

template <int NS>
__global__ __launch_bounds__(64 * CWAVES) void render_cull_kernel(RenderArgs A0, int* __restrict__ counts,
                                                                  int* __restrict__ slots) {
    extern __shared__ float smem[];
    const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // uniform per wavefront
    const int64_t t = (int64_t)blockIdx.x * CWAVES + wave;
    if (t >= A0.ntask) return;
    // a task culls cull_group consecutive chunks of one ray; ray-major, so neighbouring
    // wavefronts share the ray's depths and constants
    RenderArgs A = A0;
    const int G = A0.cull_group;
    int64_t ray;
    int cg;
    if (A0.scenes) {  // a batch: tasks cover each scene's rays x its own chunk groups
        const int* rec = A0.scenes + find_scene(A0.scenes, A0.n_scenes, 8, __builtin_amdgcn_readfirstlane((int)t)) *
                                         SCENE_REC;
        scene_strokes(A, rec);
        if (A.nch == 0) return;
        const int ngr = (A.nch + G - 1) / G;
        const int local = (int)(t - rec[8]);
        ray = rec[0] + local / ngr;
        cg = local - (local / ngr) * ngr;
    } else {
        const int ngr = (A0.nch_stride + G - 1) / G;
        ray = t / ngr;
        cg = (int)(t - ray * ngr);
    }
    WaveMem w;
    w.ts = smem + wave * cull_wave_floats(NS);
    int* ilo = (int*)(w.ts + 64 * NS);  // [CBATCH][64]
    int* ihi = ilo + CBATCH * 64;
    int* ist = ihi + CBATCH * 64;
    int* src = ist + CBATCH * 64;       // [64]
    int* cand = src + 64;               // [128]
    RayConst rc = load_ray(A, ray);
    float ts[NS];
    bool valid[NS];
    const bool sorted = stage_depths<NS>(A, ray, lane, w, ts, valid);
    if (!sorted) unsorted_range<NS>(A, rc, ts, valid);
    const int ch_end = min(A.nch, (cg + 1) * G);
    for (int cb = cg * G; cb < ch_end; cb += CBATCH) {
        const int nb = min(CBATCH, ch_end - cb);
        // (1) per chunk: cull every stroke; for hull-hit curves test each piece's chord box,
        //     one lane per piece, and keep the mask of passing pieces
        for (int j = 0; j < nb; j++) {
            const int base0 = (cb + j) * 64;
            const int k = A.S - 1 - (base0 + lane);
            float ta = 0.f, tb = -1.f;
            bool curve = false;
            const bool hit = k >= 0 && cull_stroke(A.table + (int64_t)k * REC, rc.oc, rc.dc, rc.di, rc.m, A.isr,
                                                   rc.tlo, rc.thi, ta, tb, curve);
            const unsigned long long cm = __ballot(hit && curve);
            int pmask = 0;
            if (cm) {
                const int crank = __popcll(cm & ((1ull << lane) - 1ull));
                if (hit && curve) src[crank] = lane;
                wave_sync();
                const int ncv = __popcll(cm);
                const int kb = A.S - 1 - base0;
                for (int q0 = 0; q0 < ncv; q0 += 64 / NPIECE) {
                    const int cs = q0 + lane / NPIECE;
                    const int piece = lane % NPIECE;
                    bool bh = false;
                    if (cs < ncv) {
                        const float* rec = A.table + (int64_t)(kb - src[cs]) * REC;
                        const float e = rec[R_RMAX] + rc.m + rec[R_DEV + piece];
                        const V3 a = ld3(rec + R_PTS + 3 * piece), b = ld3(rec + R_PTS + 3 * piece + 3);
                        float p0 = rc.tlo, p1 = rc.thi;
                        slab_inv(rc.oc.x, rc.di.x, fminf(a.x, b.x) - e, fmaxf(a.x, b.x) + e, p0, p1);
                        slab_inv(rc.oc.y, rc.di.y, fminf(a.y, b.y) - e, fmaxf(a.y, b.y) + e, p0, p1);
                        slab_inv(rc.oc.z, rc.di.z, fminf(a.z, b.z) - e, fmaxf(a.z, b.z) + e, p0, p1);
                        bh = p0 <= p1;
                    }
                    const unsigned long long bm = __ballot(bh);
                    if (cs < ncv && piece == 0) cand[cs] = (int)((bm >> (lane & ~(NPIECE - 1))) & 0xffffull);
                }
                wave_sync();
                if (hit && curve) pmask = cand[crank];
                wave_sync();
            }
            const bool live = hit && (!curve || pmask != 0);
            ilo[j * 64 + lane] = curve ? ord_int(1e30f) : ord_int(ta);
            ihi[j * 64 + lane] = curve ? ord_int(-1e30f) : ord_int(tb);
            ist[j * 64 + lane] = (live ? 1 : 0) | (curve ? 2 : 0) | (pmask << 16);
        }
        wave_sync();
        // (2) the passing pieces of all chunks of the batch against their chord capsules, 64
        //     at a time; each hit capsule widens its curve's interval (LDS min / max)
        int nc = 0;  // pending candidates, < 64 between rounds
        for (int j = 0; j <= nb; j++) {
            unsigned pm = 0u;
            int np = 0, incl = 0, tot = 0;
            if (j < nb) {
                const int st = ist[j * 64 + lane];
                pm = (st & 1) ? ((unsigned)st >> 16) : 0u;
                np = __popc(pm);
                incl = wave_incl_scan(np, lane);
                tot = lane_i(incl, 63);
            }
            int done = 0;
            do {
                // append up to the free buffer space of this chunk's candidates
                const int take = min(tot - done, 128 - nc);
                int idx = incl - np;
                for (unsigned m = pm; m; m &= m - 1, idx++)
                    if (idx >= done && idx < done + take)
                        cand[nc + idx - done] = (j << 10) | (lane << 4) | __builtin_ctz(m);
                nc += take;
                done += take;
                wave_sync();
                // capsule rounds: full ones while filling, the rest after the last chunk
                while (nc >= 64 || (j == nb && nc > 0)) {
                    const int nr = min(nc, 64);
                    if (lane < nr) {
                        const int cd = cand[lane];
                        const int cj = cd >> 10, cl = (cd >> 4) & 63, pc = cd & 15;
                        const float* rec = A.table + (int64_t)(A.S - 1 - ((cb + cj) * 64 + cl)) * REC;
                        float p0, p1;
                        if (capsule_interval(rc.oc, rc.dc, rc.idd, ld3(rec + R_PTS + 3 * pc),
                                             ld3(rec + R_PTS + 3 * pc + 3), rec[R_RMAX] + rc.m + rec[R_DEV + pc], p0,
                                             p1)) {
                            p0 = fmaxf(p0, rc.tlo);
                            p1 = fminf(p1, rc.thi);
                            if (p0 <= p1) {
                                atomicMin(&ilo[cj * 64 + cl], ord_int(p0));
                                atomicMax(&ihi[cj * 64 + cl], ord_int(p1));
                            }
                        }
                    }
                    const int rest = nc - nr;
                    const int mv = lane < rest ? cand[64 + lane] : 0;
                    wave_sync();
                    if (lane < rest) cand[lane] = mv;
                    nc = rest;
                    wave_sync();
                }
            } while (done < tot);
        }
        wave_sync();
        // (3) per chunk: the sample ranges, then the compacted hits
        for (int j = 0; j < nb; j++) {
            const int ch = cb + j;
            const int64_t slot = ray * A0.nch_stride + ch;
            const int k = A.S - 1 - (ch * 64 + lane);
            const int st = ist[j * 64 + lane];
            bool hit = (st & 1) != 0;
            int i0 = 0, cnt = 0, cls = 2;
            if (hit) {
                float ta = ord_float(ilo[j * 64 + lane]), tb = ord_float(ihi[j * 64 + lane]);
                if (st & 2) {  // curve: the hull of its hit capsules
                    hit = ta <= tb;
                    if (hit) pad_interval(ta, tb, ta, tb);
                }
                if (hit) {
                    if (sorted) {
                        i0 = count_below<NS, false>(w.ts, ta);
                        cnt = count_below<NS, true>(w.ts, tb) - i0;
                    } else {
                        cnt = A.n;
                    }
                    hit = cnt > 0;
                    cls = stroke_class(A.table + (int64_t)k * REC, (st & 2) != 0);
                }
            }
            const unsigned long long mask = __ballot(hit);
            const int incl = wave_incl_scan(hit ? cnt : 0, lane);
            if (hit) slots[slot * 64 + __popcll(mask & ((1ull << lane) - 1ull))] = pack_hit(lane, i0, cnt, cls);
            const int tot = lane_i(incl, 63);
            if (lane == 0) counts[slot] = __popcll(mask) | (tot << 8);
        }
        wave_sync();
    }
}

// ---------------------------------------------------------------------------
// Grid-assisted culling (one scene of at most GRID_MAX_NCH chunks).  A uniform
// grid over the strokes' culling bounds holds, per cell and 64-stroke chunk, the
// mask of the strokes whose bound (for a curve: one of its pieces' chord boxes),
// grown by the largest ray margin, overlaps the cell.  Every ray ORs the masks of
// the cells its segment crosses.  A stroke outside its ray's mask fails the
// culling tests of render_cull_kernel for that ray, so render_cull_grid_kernel
// runs those tests on the (ray, stroke) pairs of the mask only, 64 to a
// wavefront, and writes the same hits.
// ---------------------------------------------------------------------------
constexpr int GRID_MAX_NCH = 64;
constexpr int64_t GRID_WORDS = 1 << 19;  // cell-mask words (64-bit) at most, 4 MiB
constexpr int GRID_RES_MAX = 32;
constexpr int GRID_RAYS_PER_BLOCK = 256;  // rays (threads) per block of grid_rays_kernel
constexpr int GRID_SETUP_THREADS = 1024;
constexpr int GRID_CAP = 128;            // candidates per culling batch
constexpr int GRID_STRIDE = 64;          // chunks per culling task at most (one task per ray)
// parameter block: box low corner [3], cell size [3], inverse cell size [3],
// resolution, all-candidates flag, largest ray margin, cell inflation
enum { GP_LO = 0, GP_CELL = 3, GP_ICELL = 6, GP_RES = 9, GP_ALL = 10, GP_M = 11, GP_DELTA = 12, GP_COUNT = 16 };

static int grid_res(int nch) {
    int r = GRID_RES_MAX;
    while (r > 4 && (int64_t)r * r * r * nch > GRID_WORDS) r--;
    return r;
}

__device__ __forceinline__ bool finite3(V3 a) { return isfinite(a.x) && isfinite(a.y) && isfinite(a.z); }

// Axis-aligned bound of the region in which stroke k (piece `piece` of a curve) can
// pass the culling tests of a ray whose margin is at most m: the unit-space box of
// cull_stroke mapped back by R diag(s), or the chord box of the piece test.
// Returns false for a solid's piece > 0.
__device__ __forceinline__ bool grid_item_box(const float* __restrict__ rec, int piece, float m, int isr, V3& lo,
                                              V3& hi) {
    if (ibits(rec[R_CURVE]) != 0) {
        const float e = rec[R_RMAX] + m + rec[R_DEV + piece];
        const V3 a = ld3(rec + R_PTS + 3 * piece), b = ld3(rec + R_PTS + 3 * piece + 3);
        lo = v3(fminf(a.x, b.x) - e, fminf(a.y, b.y) - e, fminf(a.z, b.z) - e);
        hi = v3(fmaxf(a.x, b.x) + e, fmaxf(a.y, b.y) + e, fmaxf(a.z, b.z) + e);
        return true;
    }
    if (piece) return false;
    const float mu = isr ? m / rec[R_MINS] : m;
    float e[3], h[3];
#pragma unroll
    for (int a = 0; a < 3; a++) e[a] = (rec[R_E0 + a] + rec[R_E1 + a] * mu) * rec[R_SC + a];
#pragma unroll
    for (int i = 0; i < 3; i++)
        h[i] = fabsf(rec[R_ROT + 3 * i]) * e[0] + fabsf(rec[R_ROT + 3 * i + 1]) * e[1] +
               fabsf(rec[R_ROT + 3 * i + 2]) * e[2];
    const V3 tr = ld3(rec + R_TR);
    lo = v3(tr.x - h[0], tr.y - h[1], tr.z - h[2]);
    hi = v3(tr.x + h[0], tr.y + h[1], tr.z + h[2]);
    return true;
}

// Per ray: the segment [tlo, thi] and margin m that the culling pass derives from sorted
// sample depths (load_ray: the first and last interval midpoints), and the block's
// largest margin; the whole launch also clears the cell masks.  A ray whose depths are
// not sorted ignores its mask (render_cull_grid_kernel tests it against every stroke),
// so its segment here need not bound its samples.
__global__ __launch_bounds__(256) void grid_rays_kernel(RenderArgs A, float* __restrict__ seg, float* __restrict__ part,
                                                        unsigned long long* __restrict__ cells, int64_t ncell_words) {
    const int64_t gt = (int64_t)blockIdx.x * blockDim.x + threadIdx.x, gn = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = gt; i < ncell_words; i += gn) cells[i] = 0ull;
    __shared__ float bm[4];
    const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // uniform per wavefront
    float m = 0.f;
    if (gt < A.R) {
        const float* td = A.tdist + gt * (A.n + 1);
        const float tlo = 0.5f * (td[0] + td[1]), thi = 0.5f * (td[A.n - 1] + td[A.n]);
        const float radii = A.radii[gt];
        const float rmax = A.cone ? radii * fmaxf(fabsf(tlo), fabsf(thi)) : radii;
        m = A.delta > 0.f ? A.delta * rmax * A.tail : 1e-6f;
        seg[gt * 3] = tlo;
        seg[gt * 3 + 1] = thi;
        seg[gt * 3 + 2] = m;
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off, 64));
    if (lane == 0) bm[wave] = m;
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.x] = fmaxf(fmaxf(bm[0], bm[1]), fmaxf(bm[2], bm[3]));
}

// One block: the largest ray margin, the box of every stroke's (piece's) bound grown
// by it, and the grid over their union (an all-candidates flag when a bound or a
// margin is not finite).
__global__ __launch_bounds__(GRID_SETUP_THREADS) void grid_setup_kernel(RenderArgs A, const float* __restrict__ part,
                                                                        int nparts, int res, float* __restrict__ gp) {
    __shared__ float red[GRID_SETUP_THREADS / 64][8];
    const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // uniform per wavefront
    float m = 0.f;
    bool bad = false;
#pragma unroll 4
    for (int i = threadIdx.x; i < nparts; i += GRID_SETUP_THREADS) {
        const float v = part[i];
        if (!isfinite(v)) bad = true;
        m = fmaxf(m, v);
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off, 64));
    if (lane == 0) red[wave][0] = m;
    __syncthreads();
    m = 0.f;
    for (int q = 0; q < GRID_SETUP_THREADS / 64; q++) m = fmaxf(m, red[q][0]);
    __syncthreads();
    V3 lo = v3s(INFINITY), hi = v3s(-INFINITY);
    for (int k = threadIdx.x; k < A.S; k += GRID_SETUP_THREADS) {
        const float* rec = A.table + (int64_t)k * REC;
        const int np = ibits(rec[R_CURVE]) != 0 ? NPIECE : 1;
#pragma unroll
        for (int pc = 0; pc < NPIECE; pc++) {
            if (pc < np) {
                V3 a, b;
                grid_item_box(rec, pc, m, A.isr, a, b);
                if (!finite3(a) || !finite3(b)) bad = true;
                lo = v3(fminf(lo.x, a.x), fminf(lo.y, a.y), fminf(lo.z, a.z));
                hi = v3(fmaxf(hi.x, b.x), fmaxf(hi.y, b.y), fmaxf(hi.z, b.z));
            }
        }
    }
    float v[6] = {lo.x, lo.y, lo.z, -hi.x, -hi.y, -hi.z};
#pragma unroll
    for (int c = 0; c < 6; c++) {
        float x = v[c];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) x = fminf(x, __shfl_xor(x, off, 64));
        v[c] = x;
    }
    // the flag is a max: reduce its negation as a min
    float f = bad ? -1.f : 0.f;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) f = fminf(f, __shfl_xor(f, off, 64));
    if (lane == 0) {
        for (int c = 0; c < 6; c++) red[wave][c] = v[c];
        red[wave][6] = f;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        float r[7] = {INFINITY, INFINITY, INFINITY, INFINITY, INFINITY, INFINITY, 0.f};
        for (int q = 0; q < GRID_SETUP_THREADS / 64; q++) {
            for (int c = 0; c < 6; c++) r[c] = fminf(r[c], red[q][c]);
            r[6] = fminf(r[6], red[q][6]);
        }
        const V3 blo = v3(r[0], r[1], r[2]), bhi = v3(-r[3], -r[4], -r[5]);
        bool all = r[6] < 0.f || !isfinite(m) || !finite3(blo) || !finite3(bhi);
        const float ext = fmaxf(bhi.x - blo.x, fmaxf(bhi.y - blo.y, bhi.z - blo.z));
        const float mag = fmaxf(fmaxf(fabsf(blo.x), fabsf(blo.y)), fmaxf(fmaxf(fabsf(blo.z), fabsf(bhi.x)),
                                                                     fmaxf(fabsf(bhi.y), fabsf(bhi.z))));
        // cells are grown by 1 % of a cell (and a relative 1e-6) against rounding in
        // the bound and in the walk
        const float delta = 0.01f * ext / (float)res + 1e-6f * (1.f + mag);
        if (!isfinite(delta)) all = true;
        const float lo3[3] = {blo.x - delta, blo.y - delta, blo.z - delta};
        const float hi3[3] = {bhi.x + delta, bhi.y + delta, bhi.z + delta};
        for (int a = 0; a < 3; a++) {
            const float cell = (hi3[a] - lo3[a]) / (float)res;
            gp[GP_LO + a] = lo3[a];
            gp[GP_CELL + a] = cell;
            gp[GP_ICELL + a] = 1.f / cell;
            if (!(cell > 0.f) || !isfinite(1.f / cell)) all = true;
        }
        gp[GP_RES] = (float)res;
        gp[GP_ALL] = all ? 1.f : 0.f;
        gp[GP_M] = m;
        gp[GP_DELTA] = delta;
    }
}

__device__ __forceinline__ int grid_cell_of(float x, float lo, float icell, int res) {
    const float c = floorf((x - lo) * icell);
    return (int)fminf(fmaxf(c, 0.f), (float)(res - 1));
}

// One wavefront per (stroke, piece): sets the stroke's bit in the masks of the cells
// its grown bound overlaps.
__global__ __launch_bounds__(256) void grid_insert_kernel(RenderArgs A, const float* __restrict__ gp,
                                                          unsigned long long* __restrict__ cells) {
    const int lane = threadIdx.x & 63;
    const int64_t it = (int64_t)blockIdx.x * 4 + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    if (it >= (int64_t)A.S * NPIECE || gp[GP_ALL] != 0.f) return;
    const int k = (int)(it / NPIECE), piece = (int)(it % NPIECE);
    V3 lo, hi;
    if (!grid_item_box(A.table + (int64_t)k * REC, piece, gp[GP_M], A.isr, lo, hi)) return;
    const int res = (int)gp[GP_RES];
    const float d = gp[GP_DELTA];
    const int ax = grid_cell_of(lo.x - d, gp[GP_LO], gp[GP_ICELL], res);
    const int ay = grid_cell_of(lo.y - d, gp[GP_LO + 1], gp[GP_ICELL + 1], res);
    const int az = grid_cell_of(lo.z - d, gp[GP_LO + 2], gp[GP_ICELL + 2], res);
    const int nx = grid_cell_of(hi.x + d, gp[GP_LO], gp[GP_ICELL], res) - ax + 1;
    const int ny = grid_cell_of(hi.y + d, gp[GP_LO + 1], gp[GP_ICELL + 1], res) - ay + 1;
    const int nz = grid_cell_of(hi.z + d, gp[GP_LO + 2], gp[GP_ICELL + 2], res) - az + 1;
    const int pos = A.S - 1 - k;
    const int chunk = pos >> 6;
    const unsigned long long bit = 1ull << (pos & 63);
    const int total = nx * ny * nz;
    for (int q = lane; q < total; q += 64) {
        const int z = q % nz, y = (q / nz) % ny, x = q / (nz * ny);
        const int64_t cell = ((int64_t)(ax + x) * res + (ay + y)) * res + (az + z);
        atomicOr(&cells[cell * A.nch + chunk], bit);
    }
}

// One wavefront per (GRID_WALK_RAYS rays, GRID_WALK_CH consecutive chunks).  First the
// lanes are (ray, piece of its segment): each walks the cells its piece crosses (cell
// by cell, each crossing depth recomputed from the cell index) and lists them in LDS.
// Then the lanes are (listed cell, chunk): GRID_WALK_CH lanes read one cell's chunk
// masks together, and the ray's masks are the OR over its cells.  A ray with more
// cells than the list holds, or with a non-finite segment, takes every stroke.
constexpr int GRID_WALK_CH = 8;
constexpr int GRID_WALK_RAYS = 8;
constexpr int GRID_WALK_PIECES = 8;
constexpr int GRID_WALK_LIST = 128;
constexpr int GRID_WALK_WAVES = 4;

__global__ __launch_bounds__(64 * GRID_WALK_WAVES) void grid_walk_kernel(RenderArgs A, const float* __restrict__ gp,
                                                                         const float* __restrict__ seg,
                                                                         const unsigned long long* __restrict__ cells,
                                                                         unsigned long long* __restrict__ rmask) {
    static_assert(GRID_WALK_RAYS * GRID_WALK_PIECES == 64 && GRID_WALK_CH * 8 == 64, "lane maps");
    __shared__ int lists[GRID_WALK_WAVES][GRID_WALK_RAYS][GRID_WALK_LIST];
    __shared__ int nlist[GRID_WALK_WAVES][GRID_WALK_RAYS];
    const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // uniform per wavefront
    const int ngrp = (A.nch + GRID_WALK_CH - 1) / GRID_WALK_CH;
    const int64_t w = (int64_t)blockIdx.x * GRID_WALK_WAVES + wave;
    const int64_t nblk = (A.R + GRID_WALK_RAYS - 1) / GRID_WALK_RAYS;
    if (w >= nblk * ngrp) return;  // wave-uniform
    const int64_t rb = w / ngrp;
    const int j0 = (int)(w - rb * ngrp) * GRID_WALK_CH;
    int* cnt = nlist[wave];
    if (lane < GRID_WALK_RAYS) cnt[lane] = 0;
    wave_sync();
    // ---- (1) lanes = (ray r, piece q)
    {
        const int r = lane / GRID_WALK_PIECES, q = lane % GRID_WALK_PIECES;
        const int64_t ray = rb * GRID_WALK_RAYS + r;
        const bool on = ray < A.R;
        const int64_t rr = on ? ray : 0;
        const float tlo = seg[rr * 3], thi = seg[rr * 3 + 1], m = seg[rr * 3 + 2];
        V3 o = ld3(A.origins + 3 * rr), d = ld3(A.dirs + 3 * rr);
        if (A.coord_div != 1.f) {
            o = v3(o.x / A.coord_div, o.y / A.coord_div, o.z / A.coord_div);
            d = v3(d.x / A.coord_div, d.y / A.coord_div, d.z / A.coord_div);
        }
        bool done = !on;
        if (on && (gp[GP_ALL] != 0.f || !isfinite(tlo) || !isfinite(thi) || !isfinite(m) || !finite3(o) ||
                   !finite3(d))) {
            if (q == 0) atomicAdd(&cnt[r], GRID_WALK_LIST + 1);  // every stroke
            done = true;
        }
        const int res = (int)gp[GP_RES];
        const float lo[3] = {gp[GP_LO], gp[GP_LO + 1], gp[GP_LO + 2]};
        const float cs[3] = {gp[GP_CELL], gp[GP_CELL + 1], gp[GP_CELL + 2]};
        const float ic[3] = {gp[GP_ICELL], gp[GP_ICELL + 1], gp[GP_ICELL + 2]};
        const float oo[3] = {o.x, o.y, o.z}, dd[3] = {d.x, d.y, d.z};
        // crossing depths use the reciprocal direction (within an ulp of the quotient; the
        // cells are grown by GP_DELTA against that); an axis the ray moves along by less
        // than 1e-30 per unit depth is never crossed
        float id[3];
#pragma unroll
        for (int a = 0; a < 3; a++) id[a] = fabsf(dd[a]) > 1e-30f ? 1.f / dd[a] : 0.f;
        // the segment, padded, clipped to the grid box (with the reciprocal too: the
        // strokes' bounds lie GP_DELTA inside the box)
        const float pad = 1e-4f * (thi - tlo) + 1e-5f * (fabsf(tlo) + fabsf(thi)) + 1e-6f;
        float t0 = tlo - pad, t1 = thi + pad;
#pragma unroll
        for (int a = 0; a < 3; a++) {
            const float blo = lo[a], bhi = lo[a] + cs[a] * (float)res;
            if (id[a] != 0.f) {
                const float u = (blo - oo[a]) * id[a], v = (bhi - oo[a]) * id[a];
                t0 = fmaxf(t0, fminf(u, v));
                t1 = fminf(t1, fmaxf(u, v));
            } else if (oo[a] < blo || oo[a] > bhi) {
                t1 = -INFINITY;
            }
        }
        // this lane's piece of [t0, t1], overlapping its neighbours'
        if (!(t0 <= t1)) done = true;
        if (!done) {
            const float len = t1 - t0, ov = 1e-3f * len + 1e-6f * (1.f + fabsf(t0) + fabsf(t1));
            const float ta = fmaxf(t0, t0 + len * (float)q / GRID_WALK_PIECES - ov);
            const float tb = fminf(t1, t0 + len * (float)(q + 1) / GRID_WALK_PIECES + ov);
            t0 = ta;
            t1 = tb;
        }
        int c[3] = {0, 0, 0}, st[3] = {0, 0, 0};
        float nx[3] = {INFINITY, INFINITY, INFINITY};
        if (!done) {
#pragma unroll
            for (int a = 0; a < 3; a++) {
                c[a] = grid_cell_of(oo[a] + t0 * dd[a], lo[a], ic[a], res);
                st[a] = id[a] > 0.f ? 1 : (id[a] < 0.f ? -1 : 0);
                nx[a] = st[a] ? (lo[a] + (float)(c[a] + (st[a] > 0)) * cs[a] - oo[a]) * id[a] : INFINITY;
            }
        }
        const int max_steps = 3 * res + 3;
        for (int steps = 0; !done; steps++) {
            const int pos = atomicAdd(&cnt[r], 1);
            if (pos < GRID_WALK_LIST) lists[wave][r][pos] = (c[0] * res + c[1]) * res + c[2];
            // step along the axis crossed first (branch-free over the three axes)
            const bool x0 = nx[0] <= nx[1] && nx[0] <= nx[2];
            const bool x1 = !x0 && nx[1] <= nx[2];
            const float tn = x0 ? nx[0] : (x1 ? nx[1] : nx[2]);
            if (pos >= GRID_WALK_LIST || !(tn <= t1) || steps + 1 >= max_steps) {
                done = true;
            } else {
#pragma unroll
                for (int a = 0; a < 3; a++) {
                    const bool sel = a == 0 ? x0 : (a == 1 ? x1 : (!x0 && !x1));
                    if (sel) {
                        c[a] += st[a];
                        nx[a] = (lo[a] + (float)(c[a] + (st[a] > 0)) * cs[a] - oo[a]) * id[a];
                    }
                }
                if (c[0] < 0 || c[0] >= res || c[1] < 0 || c[1] >= res || c[2] < 0 || c[2] >= res) done = true;
            }
        }
    }
    wave_sync();
    // ---- (2) lanes = (listed cell u, chunk cc), one ray at a time
    const int u = lane / GRID_WALK_CH, cc = lane % GRID_WALK_CH;
    const bool has = j0 + cc < A.nch;
    for (int r = 0; r < GRID_WALK_RAYS; r++) {
        const int64_t ray = rb * GRID_WALK_RAYS + r;
        if (ray >= A.R) break;  // wave-uniform
        const int n = cnt[r];
        unsigned long long acc = 0ull;
        if (n > GRID_WALK_LIST) {
            acc = ~0ull;
        } else {
            for (int k0 = 0; k0 < n; k0 += 8) {
                const int k = k0 + u;
                if (k < n && has) acc |= cells[(int64_t)lists[wave][r][k] * A.nch + j0 + cc];
            }
        }
#pragma unroll
        for (int off = GRID_WALK_CH; off < 64; off <<= 1) {
            const unsigned l32 = __shfl_xor((unsigned)acc, off, 64), h32 = __shfl_xor((unsigned)(acc >> 32), off, 64);
            acc |= ((unsigned long long)h32 << 32) | l32;
        }
        if (u == 0 && has) {
            const int nvalid = min(64, A.S - 64 * (j0 + cc));
            rmask[ray * A.nch + j0 + cc] = acc & (nvalid >= 64 ? ~0ull : ((1ull << nvalid) - 1ull));
        }
    }
}

// per wave: depths [64 NS]; per batch candidate: interval lo / hi (order-preserving
// ints), state (hit | curve << 1 | passing pieces << 16) and (chunk << 6 | lane); curve
// -> candidate map [64]; capsule candidates [128]; per task chunk: first candidate,
// hits and pairs written
__host__ __device__ constexpr int cullg_wave_floats(int ns) {
    return 64 * ns + GRID_CAP * 4 + 64 + 128 + 3 * (GRID_STRIDE + 1);
}

template <int NS>
__global__ __launch_bounds__(64 * CWAVES) void render_cull_grid_kernel(RenderArgs A,
                                                                       const unsigned long long* __restrict__ rmask,
                                                                       int* __restrict__ counts,
                                                                       int* __restrict__ slots) {
    extern __shared__ float smem[];
    const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // uniform per wavefront
    const int64_t ray = (int64_t)blockIdx.x * CWAVES + wave;
    if (ray >= A.R) return;
    WaveMem w;
    w.ts = smem + wave * cullg_wave_floats(NS);
    int* ilo = (int*)(w.ts + 64 * NS);  // [GRID_CAP]
    int* ihi = ilo + GRID_CAP;
    int* ist = ihi + GRID_CAP;
    int* cl = ist + GRID_CAP;           // candidate -> chunk << 6 | lane
    int* src = cl + GRID_CAP;           // [64]
    int* cand = src + 64;               // [128]
    int* cfirst = cand + 128;           // [GRID_STRIDE + 1] first candidate of each chunk
    int* chits = cfirst + GRID_STRIDE + 1;
    int* cpairs = chits + GRID_STRIDE + 1;
    RayConst rc = load_ray(A, ray);
    float ts[NS];
    bool valid[NS];
    const bool sorted = stage_depths<NS>(A, ray, lane, w, ts, valid);
    if (!sorted) unsorted_range<NS>(A, rc, ts, valid);
    const int nch = A.nch;
    // lane j: the mask of chunk j (every stroke of it when the depths are not sorted) and
    // the candidates before it
    const int nvalid = min(64, A.S - 64 * lane);
    const unsigned long long mj =
        lane >= nch ? 0ull
                    : (sorted ? rmask[ray * nch + lane] : (nvalid >= 64 ? ~0ull : ((1ull << nvalid) - 1ull)));
    const int pc = __popcll(mj);
    const int pincl = wave_incl_scan(pc, lane);
    if (lane < nch) {
        cfirst[lane] = pincl - pc;
        chits[lane] = 0;
        cpairs[lane] = 0;
    }
    const int total = lane_i(pincl, 63);
    if (lane == 0) cfirst[nch] = total;
    wave_sync();
    const int64_t slot0 = ray * A.nch_stride;
    for (int jb = 0, cb = 0; jb < nch; ) {
        // the batch: chunks [jb, je) with at most GRID_CAP candidates (at least one chunk)
        const unsigned long long fit = __ballot(lane >= jb && lane < nch && pincl - cb <= GRID_CAP);
        int je = fit ? 64 - __builtin_clzll(fit) : jb + 1;
        if (je <= jb) je = jb + 1;
        const int nc = cfirst[je] - cb;
        for (int j = jb; j < je; j++) {
            const unsigned long long mm = __shfl(mj, j, 64);  // 64-bit shuffle of lane j's mask
            if ((mm >> lane) & 1ull) cl[cfirst[j] - cb + __popcll(mm & ((1ull << lane) - 1ull))] = (j << 6) | lane;
        }
        wave_sync();
        // (1) per candidate: cull the stroke; for hull-hit curves test each piece's chord
        //     box, one lane per piece, and keep the mask of passing pieces
        for (int r0 = 0; r0 < nc; r0 += 64) {
            const int q = r0 + lane;
            const bool on = q < nc;
            const int code = on ? cl[q] : 0;
            const int k = A.S - 1 - ((code >> 6) * 64 + (code & 63));
            float ta = 0.f, tb = -1.f;
            bool curve = false;
            const bool hit = on && cull_stroke(A.table + (int64_t)k * REC, rc.oc, rc.dc, rc.di, rc.m, A.isr,
                                               rc.tlo, rc.thi, ta, tb, curve);
            const unsigned long long cm = __ballot(hit && curve);
            int pmask = 0;
            if (cm) {
                const int crank = __popcll(cm & ((1ull << lane) - 1ull));
                if (hit && curve) src[crank] = k;
                wave_sync();
                const int ncv = __popcll(cm);
                for (int q0 = 0; q0 < ncv; q0 += 64 / NPIECE) {
                    const int cs = q0 + lane / NPIECE;
                    const int piece = lane % NPIECE;
                    bool bh = false;
                    if (cs < ncv) {
                        const float* rec = A.table + (int64_t)src[cs] * REC;
                        const float e = rec[R_RMAX] + rc.m + rec[R_DEV + piece];
                        const V3 a = ld3(rec + R_PTS + 3 * piece), b = ld3(rec + R_PTS + 3 * piece + 3);
                        float p0 = rc.tlo, p1 = rc.thi;
                        slab_inv(rc.oc.x, rc.di.x, fminf(a.x, b.x) - e, fmaxf(a.x, b.x) + e, p0, p1);
                        slab_inv(rc.oc.y, rc.di.y, fminf(a.y, b.y) - e, fmaxf(a.y, b.y) + e, p0, p1);
                        slab_inv(rc.oc.z, rc.di.z, fminf(a.z, b.z) - e, fmaxf(a.z, b.z) + e, p0, p1);
                        bh = p0 <= p1;
                    }
                    const unsigned long long bm = __ballot(bh);
                    if (cs < ncv && piece == 0) cand[cs] = (int)((bm >> (lane & ~(NPIECE - 1))) & 0xffffull);
                }
                wave_sync();
                if (hit && curve) pmask = cand[crank];
                wave_sync();
            }
            if (on) {
                const bool live = hit && (!curve || pmask != 0);
                ilo[q] = curve ? ord_int(1e30f) : ord_int(ta);
                ihi[q] = curve ? ord_int(-1e30f) : ord_int(tb);
                ist[q] = (live ? 1 : 0) | (curve ? 2 : 0) | (pmask << 16);
            }
        }
        wave_sync();
        // (2) the passing pieces against their chord capsules, 64 at a time; each hit
        //     capsule widens its curve's interval (LDS min / max)
        int np_c = 0;  // pending capsule candidates, < 64 between rounds
        const int nrounds = (nc + 63) / 64;
        for (int rr = 0; rr <= nrounds; rr++) {  // the last pass only flushes
            const int r0 = rr * 64;
            unsigned pm = 0u;
            int npc = 0, incl = 0, tot = 0;
            const bool more = rr < nrounds;
            if (more) {
                const int q = r0 + lane;
                const int st = q < nc ? ist[q] : 0;
                pm = (st & 1) ? ((unsigned)st >> 16) : 0u;
                npc = __popc(pm);
                incl = wave_incl_scan(npc, lane);
                tot = lane_i(incl, 63);
            }
            int done = 0;
            do {
                const int take = min(tot - done, 128 - np_c);
                int idx = incl - npc;
                for (unsigned mm = pm; mm; mm &= mm - 1, idx++)
                    if (idx >= done && idx < done + take)
                        cand[np_c + idx - done] = ((r0 + lane) << 4) | __builtin_ctz(mm);
                np_c += take;
                done += take;
                wave_sync();
                while (np_c >= 64 || (!more && np_c > 0)) {
                    const int nr = min(np_c, 64);
                    if (lane < nr) {
                        const int cd = cand[lane];
                        const int cq = cd >> 4, pcs = cd & 15;
                        const int code = cl[cq];
                        const float* rec = A.table + (int64_t)(A.S - 1 - ((code >> 6) * 64 + (code & 63))) * REC;
                        float p0, p1;
                        if (capsule_interval(rc.oc, rc.dc, rc.idd, ld3(rec + R_PTS + 3 * pcs),
                                             ld3(rec + R_PTS + 3 * pcs + 3), rec[R_RMAX] + rc.m + rec[R_DEV + pcs],
                                             p0, p1)) {
                            p0 = fmaxf(p0, rc.tlo);
                            p1 = fminf(p1, rc.thi);
                            if (p0 <= p1) {
                                atomicMin(&ilo[cq], ord_int(p0));
                                atomicMax(&ihi[cq], ord_int(p1));
                            }
                        }
                    }
                    const int rest = np_c - nr;
                    const int mv = lane < rest ? cand[64 + lane] : 0;
                    wave_sync();
                    if (lane < rest) cand[lane] = mv;
                    np_c = rest;
                    wave_sync();
                }
            } while (done < tot);
        }
        wave_sync();
        // (3) per candidate: the sample range; hits are written per chunk in lane order
        for (int r0 = 0; r0 < nc; r0 += 64) {
            const int q = r0 + lane;
            const bool on = q < nc;
            const int code = on ? cl[q] : 0;
            const int j = code >> 6, l = code & 63;
            const int k = A.S - 1 - (j * 64 + l);
            const int st = on ? ist[q] : 0;
            bool hit = (st & 1) != 0;
            int i0 = 0, cnt = 0, cls = 2;
            if (hit) {
                float ta = ord_float(ilo[q]), tb = ord_float(ihi[q]);
                if (st & 2) {
                    hit = ta <= tb;
                    if (hit) pad_interval(ta, tb, ta, tb);
                }
                if (hit) {
                    if (sorted) {
                        i0 = count_below<NS, false>(w.ts, ta);
                        cnt = count_below<NS, true>(w.ts, tb) - i0;
                    } else {
                        cnt = A.n;
                    }
                    hit = cnt > 0;
                    cls = stroke_class(A.table + (int64_t)k * REC, (st & 2) != 0);
                }
            }
            const unsigned long long hm = __ballot(hit);
            if (hit) {
                // candidates of one chunk are contiguous: rank among this round's hits of the
                // chunk, after the chunk's hits of earlier rounds
                const int s0 = max(cfirst[j] - cb - r0, 0);
                const unsigned long long below = (1ull << lane) - 1ull;
                const unsigned long long seg0 = s0 > 0 ? (1ull << s0) - 1ull : 0ull;
                const int rank = chits[j] + __popcll(hm & below & ~seg0);
                slots[(slot0 + j) * 64 + rank] = pack_hit(l, i0, cnt, cls);
            }
            wave_sync();
            if (hit) {
                atomicAdd(&chits[j], 1);
                atomicAdd(&cpairs[j], cnt);
            }
            wave_sync();
        }
        cb += nc;
        jb = je;
    }
    if (lane < nch) counts[slot0 + lane] = chits[lane] | (cpairs[lane] << 8);
}

// ---------------------------------------------------------------------------
// forward
// ---------------------------------------------------------------------------
// PRE: the hits come from the culling pass (render_cull_kernel); otherwise the
// forward culls chunk by chunk itself.
template <int NS, int CID, bool PRE>
__device__ __forceinline__ void render_fwd_body(const RenderArgs& A0, float* __restrict__ rgb_out,
                                                float* __restrict__ acc_out, float* __restrict__ depth_out,
                                                float* __restrict__ state,
                                                unsigned long long* __restrict__ counters) {
    extern __shared__ float smem[];
    constexpr int NPF = CID == 1 ? 4 : 1;  // alpha (+ colour for gradient_rgb)
    const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // uniform per wavefront
    const int64_t ray = (int64_t)blockIdx.x * FWAVES + wave;
    if (ray >= A0.R) return;
    RenderArgs A = A0;
    if (A0.scenes) {
        const int sc = find_scene(A0.scenes, A0.n_scenes, 0, __builtin_amdgcn_readfirstlane((int)ray));
        scene_strokes(A, A0.scenes + sc * SCENE_REC);
    }
    const WaveMem w = wave_mem(smem + wave * wave_mem_floats(NS, NPF, 0), NS, NPF);
    RayConst rc = load_ray(A, ray);
    float ts[NS];
    bool valid[NS];
    const bool sorted = stage_depths<NS>(A, ray, lane, w, ts, valid);
    if (!sorted) unsorted_range<NS>(A, rc, ts, valid);
    // T: transmittance past the strokes composited so far; U = 1 - T kept as the sum of the
    // weights (no cancellation when T ~ 1, where the normalisation 1 / (1e-6 + 1 - T) is steep)
    float T[NS], U[NS], Dn[NS];
    V3 C[NS];
#pragma unroll
    for (int s = 0; s < NS; s++) { T[s] = 1.f; U[s] = 0.f; Dn[s] = 0.f; C[s] = v3s(0.f); }
    // cursors into this ray's region of the kept hits / pair values; once a group of
    // chunks does not fit, it and every later chunk are left to the adjoint to recompute
    int hoff = 0, poff = 0;
    bool full = false;
    const int64_t rh = ray * (int64_t)A.cap_h;
    const int64_t rp = ray * (int64_t)A.cap_p * NPF;
    int nh = 0, np = 0;  // pending hits / pairs
    int gch = 0;         // first chunk of the pending group
    int cword = 0;       // culling pass: count word of chunk (ch & ~63) + lane
    int ch0 = 0;
    bool loaded = false;  // all of the ray's hits are pending already
    if constexpr (PRE) {
        // when the ray's hits fit the pending list, gather them in one pass (one
        // lane per hit; its chunk found by a binary search over the chunks' first hits)
        if (A.nch <= 64) {
            const int64_t slot0 = ray * A.nch_stride;
            cword = lane < A.nch ? A.ccounts[slot0 + lane] : 0;
            const int ncl = cword & 127;
            const int hincl = wave_incl_scan(ncl, lane);
            const int H = lane_i(hincl, 63);
            if (H <= HMAX) {
                w.xo[lane] = hincl - ncl;
                wave_sync();
                int carry = 0;
                for (int h0 = 0; h0 < H; h0 += 64) {
                    const int hi = h0 + lane;
                    int cnt = 0, c = 0, hw = 0;
                    if (hi < H) {
#pragma unroll
                        for (int step = 32; step > 0; step >>= 1)
                            if (c + step < A.nch && w.xo[c + step] <= hi) c += step;
                        hw = A.cslots[(slot0 + c) * 64 + (hi - w.xo[c])];
                        cnt = (hw >> 15) & 511;
                    }
                    const int incl = wave_incl_scan(cnt, lane);
                    if (hi < H) {
                        w.ck[hi] = A.S - 1 - (c * 64 + (hw & 63));
                        w.ci0[hi] = (hw >> 6) & 511;
                        w.ccnt[hi] = cnt;
                        w.ccls[hi] = (hw >> 24) & 3;
                        w.coff[hi] = carry + incl - cnt;
                    }
                    carry += lane_i(incl, 63);
                }
                wave_sync();
                nh = H;
                np = carry;
                ch0 = A.nch - 1;
                loaded = true;
            }
        }
    }
    for (int ch = ch0; ch < A.nch; ch++) {
        int nc = 0, tot = 0;
        if constexpr (PRE) {
          if (!loaded) {
            // hits of the culling pass, final (capsules applied)
            const int64_t slot = ray * A.nch_stride + ch;
            if ((ch & 63) == 0) cword = ch + lane < A.nch ? A.ccounts[slot + lane] : 0;
            const int cw = __builtin_amdgcn_readlane(cword, ch & 63);
            nc = cw & 127;
            tot = cw >> 8;
            if (nc) {
                const int hw = lane < nc ? A.cslots[slot * 64 + lane] : 0;
                const int cnt = lane < nc ? (hw >> 15) & 511 : 0;
                const int incl = wave_incl_scan(cnt, lane);
                if (lane < nc) {
                    w.ck[nh + lane] = A.S - 1 - (ch * 64 + (hw & 63));
                    w.ci0[nh + lane] = (hw >> 6) & 511;
                    w.ccnt[nh + lane] = cnt;
                    w.ccls[nh + lane] = (hw >> 24) & 3;
                    w.coff[nh + lane] = np + incl - cnt;
                }
                wave_sync();
            }
          }
        } else {
            nc = append_hits<NS>(A, rc, w, sorted, ch * 64, lane, nh, np, tot);
        }
        nh += nc;
        np += tot;
        if (nh <= HMAX - 64 && np < A.flush_pairs && ch + 1 < A.nch) continue;
        if constexpr (!PRE) nh = refine_pending<NS>(A, rc, w, sorted, nh, lane, np);
        int kept_np = 0;
        const int group_poff = poff;
        if (A.hdir) {
            const bool keep = !full && hoff + nh <= A.cap_h && poff + np <= A.cap_p;
            full = !keep;
            // per chunk of the group: first hit, absolute offset of its first pair, hits
            for (int cb = gch; cb <= ch; cb += 64) {
                const int cc = cb + lane;
                if (cc <= ch) {
                    const int f = first_hit_of_chunk(w, nh, A.S - 1, cc);
                    const int f1 = first_hit_of_chunk(w, nh, A.S - 1, cc + 1);
                    int* e = A.hdir + (ray * A.nch_stride + cc) * 3;
                    e[0] = keep ? hoff + f : -1;
                    e[1] = poff + (f < nh ? w.coff[f] : np);
                    e[2] = f1 - f;
                }
            }
            if (keep) {
                for (int c = lane; c < nh; c += 64) {
                    int* h = A.hbuf + (rh + hoff + c) * 4;
                    h[0] = w.ck[c] | ((w.ccls[c] & 3) << 28);  // stroke | class << 28
                    h[1] = w.ci0[c];
                    h[2] = w.ccnt[c];
                    h[3] = poff + w.coff[c];  // absolute pair offset in the ray's region
                }
                hoff += nh;
                poff += np;
                kept_np = np;
            }
        }
        gch = ch + 1;
        // evaluate and composite the pending hits, window by window
        for (int c0 = 0; c0 < nh;) {
            const int wbase = w.coff[c0];
            const int c1 = window_end(w, c0, nh, wbase, lane);
            const int wp = w.coff[c1 - 1] + w.ccnt[c1 - 1] - wbase;
            // alpha of every (sample, stroke) pair of the window, one pair per lane, in class
            // order (a 64-pair group spans at most a class boundary or two)
            const int hc = c0 + lane;
            const int hcls = hc < c1 ? (w.ccls[hc] & 3) : -1;
            {
              const int nx = class_sort_hits(w, hc, hcls, lane);
              const int xt = wp;
              for (int p0 = 0; p0 < xt; p0 += 64) {
                const int pp = p0 + lane;
                if (pp < xt) {
                    const int xr = class_find(w, nx, pp);
                    const int c = w.xh[xr];
                    const int j = pp - w.xo[xr];
                    const int p = w.coff[c] - wbase + j;  // position in the window
                    const int k = w.ck[c];
                    const int i = w.ci0[c] + j;
                    V3 x;
                    float r;
                    sample_point(A, rc, w.ts[i], x, r);
                    const float* rec = A.table + (int64_t)k * REC;
                    const Pair q = eval_pair(rec, ibits(rec[R_BASE]), ibits(rec[R_CURVE]) != 0, x, r, A);
                    const bool keep = wbase + p < kept_np;
                    w.pb[p] = q.alpha;
                    if (keep) A.pbuf[rp + group_poff + wbase + p] = q.alpha;
                    if (CID == 1) {
                        const V3 col = color_eval_t<1>(q.p, rc.vd, A.cparams + (int64_t)k * A.Q);
                        w.pb[PBUF + p] = col.x;
                        w.pb[2 * PBUF + p] = col.y;
                        w.pb[3 * PBUF + p] = col.z;
                        if (keep) {
                            float* kp = A.pbuf + rp + group_poff + wbase + p;
                            kp[A.cap_p] = col.x;
                            kp[2 * (int64_t)A.cap_p] = col.y;
                            kp[3 * (int64_t)A.cap_p] = col.z;
                        }
                    }
                }
              }
            }
            wave_sync();
            // ordered over-compositing, strokes front to back
            const WindowHits<CID> wh = window_hits<CID>(A, w, c0, c1, wbase, lane, rc.vd);
            for (int c = c0; c < c1; c++) {
                const int j = c - c0;
                const int k = lane_i(wh.k, j);
                const int ci = lane_i(wh.ci, j);
                const int cn = lane_i(wh.cn, j);
                const int po = lane_i(wh.po, j);
                const float dk = lane_f(wh.dk, j);
                const V3 cu = v3(lane_f(wh.cu.x, j), lane_f(wh.cu.y, j), lane_f(wh.cu.z, j));
                int tp = 0;
#pragma unroll
                for (int s = 0; s < NS; s++) {
                    const int j = lane + 64 * s - ci;
                    bool pos = false;
                    if ((unsigned)j < (unsigned)cn) {
                        const int q = po + j;
                        const float a = w.pb[q];
                        if (a != 0.f) {
                            const V3 col = (CID == 1) ? v3(w.pb[PBUF + q], w.pb[2 * PBUF + q], w.pb[3 * PBUF + q]) : cu;
                            const float wt = a * T[s];
                            T[s] *= (1.f - a);
                            U[s] += wt;
                            Dn[s] += dk * wt;
                            C[s] = C[s] + col * wt;
                        }
                        pos = a > 0.f;
                    }
                    if (counters) tp += __popcll(__ballot(pos));
                }
                if (counters) {
                    const int base = ibits(A.table[(int64_t)k * REC + R_BASE]);
                    // counting pass only: per stroke type, pairs with alpha > 0, evaluated
                    // pairs and hits (atomics, so no registers stay live in the timed path)
                    if (lane == 0) {
                        atomicAdd(counters + base, (unsigned long long)tp);
                        atomicAdd(counters + NUM_BASE + base, (unsigned long long)cn);
                        atomicAdd(counters + 2 * NUM_BASE, 1ull);
                    }
                }
            }
            wave_sync();
            c0 = c1;
        }
        nh = np = kept_np = 0;
    }
    // save per-sample composite state and run the ray march
    const int64_t N = A.R * (int64_t)A.n;
    const float* td = A.tdist + ray * (A.n + 1);
    const float dn = sqrtf(rc.d.x * rc.d.x + rc.d.y * rc.d.y + rc.d.z * rc.d.z);
    float sd[NS], rgbn[NS][3];
#pragma unroll
    for (int s = 0; s < NS; s++) {
        const int i = lane + 64 * s;
        const float scl = 1.f / (1e-6f + U[s]);
        rgbn[s][0] = clampf(C[s].x * scl, 0.f, 1.f);
        rgbn[s][1] = clampf(C[s].y * scl, 0.f, 1.f);
        rgbn[s][2] = clampf(C[s].z * scl, 0.f, 1.f);
        if (valid[s]) {
            const int64_t si = ray * A.n + i;
            state[si] = Dn[s];
            state[N + si] = C[s].x;
            state[2 * N + si] = C[s].y;
            state[3 * N + si] = C[s].z;
            state[4 * N + si] = U[s];
            sd[s] = Dn[s] * ((td[i + 1] - td[i]) * dn);
        } else {
            sd[s] = 0.f;
        }
    }
    // exclusive prefix sum of density*delta over samples (slot-major order)
    float carry = 0.f;
    float acc = 0.f, r0 = 0.f, r1 = 0.f, r2 = 0.f, dsum = 0.f;
#pragma unroll
    for (int s = 0; s < NS; s++) {
        float v = sd[s];
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const float u = __shfl_up(v, off, 64);
            if (lane >= off) v += u;
        }
        const float excl = carry + v - sd[s];
        const float trans = expf(-excl);
        const float al = (A.opaque && lane + 64 * s == A.n - 1) ? 1.f : 1.f - expf(-sd[s]);
        const float wt = valid[s] ? al * trans : 0.f;
        carry += lane_f(v, 63);
        acc += wt;
        r0 += wt * rgbn[s][0];
        r1 += wt * rgbn[s][1];
        r2 += wt * rgbn[s][2];
        dsum += wt * ts[s];
    }
    acc = wave_sum(acc);
    r0 = wave_sum(r0);
    r1 = wave_sum(r1);
    r2 = wave_sum(r2);
    dsum = wave_sum(dsum);
    if (lane == 0) {
        const float bgw = fmaxf(1.f - acc, 0.f);
        const V3 bgc = A.bgs ? ld3(A.bgs + 3 * ray) : v3(A.bg0, A.bg1, A.bg2);
        rgb_out[3 * ray + 0] = r0 + bgw * bgc.x;
        rgb_out[3 * ray + 1] = r1 + bgw * bgc.y;
        rgb_out[3 * ray + 2] = r2 + bgw * bgc.z;
        acc_out[ray] = acc;
        float dep = dsum / fmaxf(acc, 1.1920929e-07f);
        if (isnan(dep)) dep = INFINITY;
        depth_out[ray] = fminf(fmaxf(dep, td[0]), td[A.n]);
    }
}

// ---------------------------------------------------------------------------
// eval-time extras of the ray march
// ---------------------------------------------------------------------------
// volumetric_rendering(compute_extras=True) (source/utils/render.py:88-116) from
// the forward's per-sample composite density (state row 0): the weights as the
// forward's ray march forms them (compute_alpha_weights, render.py:29-48), the
// log-distance mean, and the 5 / 50 / 95 weighted percentiles of the fenceposts
// extended by the far distance with the background weight
// (stepfun.weighted_percentile over integrate_weights, located by sorted_interp,
// stepfun.py:5-24,149-170,351-361).  For a percentile x the fenceposts split by
// cw_j <= x; the interval is (largest cw / t of the first set, smallest of the
// second), defaulting to (cw, t) = (0, t_0) and (1, t_far).  One wavefront per
// ray; out [R, 4] = distance mean, 5th percentile, median, 95th percentile.
constexpr int XWAVES = 4;

__global__ __launch_bounds__(64 * XWAVES) void render_extras_kernel(const float* __restrict__ state,
                                                                    const float* __restrict__ tdist,
                                                                    const float* __restrict__ dirs,
                                                                    const float* __restrict__ t_far, int64_t R,
                                                                    int n, int opaque, float* __restrict__ out) {
    const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // uniform per wavefront
    const int64_t ray = (int64_t)blockIdx.x * XWAVES + wave;
    if (ray >= R) return;
    const float* td = tdist + ray * (n + 1);
    const float* dens = state + ray * n;
    const V3 d = ld3(dirs + 3 * ray);
    const float dn = sqrtf(d.x * d.x + d.y * d.y + d.z * d.z);
    const float far = t_far[ray];
    const float ps[3] = {0.05f, 0.5f, 0.95f};
    float x0[3], x1[3], f0[3], f1[3];
#pragma unroll
    for (int q = 0; q < 3; q++) {
        x0[q] = 0.f;
        f0[q] = td[0];
        x1[q] = 1.f;
        f1[q] = far;
    }
    float carry = 0.f, wcarry = 0.f, acc = 0.f, lsum = 0.f;
    for (int c0 = 0; c0 < n; c0 += 64) {
        const int i = c0 + lane;
        const bool valid = i < n;
        const float ta = valid ? td[i] : 0.f, tb = valid ? td[i + 1] : 0.f;
        const float sd = valid ? dens[i] * ((tb - ta) * dn) : 0.f;
        float v = sd;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const float u = __shfl_up(v, off, 64);
            if (lane >= off) v += u;
        }
        const float trans = expf(-(carry + v - sd));
        const float al = (opaque && i == n - 1) ? 1.f : 1.f - expf(-sd);
        const float w = valid ? al * trans : 0.f;
        carry += lane_f(v, 63);
        acc += w;
        if (valid) lsum += w * logf(0.5f * (ta + tb));
        // fencepost i + 1: the integrated weight of samples 0..i (clamped to 1) at depth td[i + 1]
        float cw = w;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const float u = __shfl_up(cw, off, 64);
            if (lane >= off) cw += u;
        }
        const float cwi = fminf(wcarry + cw, 1.f);
        wcarry += lane_f(cw, 63);
        if (valid) {
#pragma unroll
            for (int q = 0; q < 3; q++) {
                if (ps[q] >= cwi) {
                    x0[q] = fmaxf(x0[q], cwi);
                    f0[q] = fmaxf(f0[q], tb);
                } else {
                    x1[q] = fminf(x1[q], cwi);
                    f1[q] = fminf(f1[q], tb);
                }
            }
        }
    }
    acc = wave_sum(acc);
    lsum = wave_sum(lsum);
#pragma unroll
    for (int q = 0; q < 3; q++) {
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            x0[q] = fmaxf(x0[q], __shfl_xor(x0[q], off, 64));
            f0[q] = fmaxf(f0[q], __shfl_xor(f0[q], off, 64));
            x1[q] = fminf(x1[q], __shfl_xor(x1[q], off, 64));
            f1[q] = fminf(f1[q], __shfl_xor(f1[q], off, 64));
        }
    }
    if (lane == 0) {
        float dm = expf(lsum / fmaxf(acc, 1.1920929e-07f));
        if (isnan(dm)) dm = INFINITY;
        out[4 * ray] = fminf(fmaxf(dm, td[0]), td[n]);
#pragma unroll
        for (int q = 0; q < 3; q++) {
            float o = (ps[q] - x0[q]) / (x1[q] - x0[q]);
            if (isnan(o)) o = 0.f;
            o = fminf(fmaxf(o, 0.f), 1.f);
            out[4 * ray + 1 + q] = f0[q] + o * (f1[q] - f0[q]);
        }
    }
}

// ---------------------------------------------------------------------------
// backward
// ---------------------------------------------------------------------------
// canonical gradient slot -> column in the [Pmax | Q | 1] accumulator row (-1: none)
// slots: 0..13 shape (solids: 0-1 base, 2-4 scale, 5-7 euler, 8-10 translation;
// curves: their 11 or 14 params), 14..14+Q-1 colour, 14+Q density
__device__ __forceinline__ int slot_column(int slot, const float* __restrict__ rec, int Pmax, int Q) {
    const bool curve = ibits(rec[R_CURVE]) != 0;
    const int nb = ibits(rec[R_NB]);
    if (slot >= 14) {
        int c = slot - 14;
        if (c < Q) return Pmax + c;
        if (c == Q) return Pmax + Q;
        return -1;
    }
    if (curve) return slot < nb ? slot : -1;
    if (slot < 2) return slot < nb ? slot : -1;
    const uint32_t flags = (uint32_t)ibits(rec[R_FLAGS]);
    if (slot < 5) {
        int off = ibits(rec[R_OFFS]);
        if (off < 0) return -1;
        if (flags & F_SS) return slot == 2 ? off : -1;
        return off + slot - 2;
    }
    if (slot < 8) {
        int off = ibits(rec[R_OFFA]);
        return off < 0 ? -1 : off + slot - 5;
    }
    if (slot < 11) {
        int off = ibits(rec[R_OFFT]);
        return off < 0 ? -1 : off + slot - 8;
    }
    return -1;
}

template <int NS, int CID>
__global__ __launch_bounds__(64 * FWAVES) __attribute__((amdgpu_waves_per_eu(3, 8))) void render_fwd_kernel(
    RenderArgs A0, float* __restrict__ rgb_out, float* __restrict__ acc_out, float* __restrict__ depth_out,
    float* __restrict__ state, unsigned long long* __restrict__ counters) {
    render_fwd_body<NS, CID, false>(A0, rgb_out, acc_out, depth_out, state, counters);
}

template <int NS, int CID>
__global__ __launch_bounds__(64 * FWAVES) __attribute__((amdgpu_waves_per_eu(4, 8))) void render_fwd_pre_kernel(
    RenderArgs A0, float* __restrict__ rgb_out, float* __restrict__ acc_out, float* __restrict__ depth_out,
    float* __restrict__ state, unsigned long long* __restrict__ counters) {
    render_fwd_body<NS, CID, true>(A0, rgb_out, acc_out, depth_out, state, counters);
}

// Parameter-gradient slots G (slot_column layout, colour from 14, density 14 + Q) of
// one (sample, stroke) pair: stroke k, sample position x and radius r, view direction
// vd, and the pair's upstream dL/dalpha, dL/d(density param) and dL/d(colour).
template <int CID>
__device__ __forceinline__ void pair_grad(const RenderArgs& A, int k, V3 x, float r, V3 vd, float ga, float gdw,
                                          V3 gcol, float (&G)[15 + num_color_params(CID)]) {
    constexpr int NQ = num_color_params(CID);
    const float* rec = A.table + (int64_t)k * REC;
    const int base = ibits(rec[R_BASE]);
    const bool curve = ibits(rec[R_CURVE]) != 0;
    const uint32_t flags = (uint32_t)ibits(rec[R_FLAGS]);
    const float* cp = A.cparams + (int64_t)k * A.Q;
    const float* prm = curve ? rec + R_CP : rec + R_BP;
    if (curve && A.exact) {
        // curves carry no transform: one closest-point solve gives the value (hence alpha
        // and dL/dSDF) and the parameter adjoint
        const ExactTransfer tr{ga, (A.laplace ? 2.f : 0.5f) / (A.delta * r), A.delta, A.laplace != 0};
        float sdf, gs;
        curve_value_grad(base, x, prm, tr, G, sdf, gs);
        color_grad_t<CID>(x, vd, cp, gcol, G + 14, true);
        G[14 + NQ] += gdw;
        return;
    }
    const Pair q = eval_pair(rec, base, curve, x, r, A);
    V3 gpos = color_grad_t<CID>(q.p, vd, cp, gcol, G + 14, A.exact != 0);
    G[14 + NQ] += gdw;
    const float gs = ga * dalpha_dsdf(q.sdf, q.alpha, q.sc, A.delta, A.laplace != 0, A.exact != 0);
    const V3 dsdp = A.exact ? grad_ex(base, q.p, prm, gs, G) : grad_ref(base, q.p, prm, gs, G);
    gpos = gpos + dsdp * gs;
    if (!curve) {
        float gs1 = 0.f;
        V3 gs3 = v3s(0.f), gang = v3s(0.f), gtr = v3s(0.f);
        Frame F;
        F.trans = ld3(rec + R_TR);
        F.angle = ld3(rec + R_ANG);
        F.sz = rec[R_SZ];
        F.cz = rec[R_CZ];
        F.scale = ld3(rec + R_SC);
#pragma unroll
        for (int ii = 0; ii < 3; ii++)
#pragma unroll
            for (int jj = 0; jj < 3; jj++) F.R.r[ii][jj] = rec[R_ROT + 3 * ii + jj];
        if (A.exact && A.isr && (flags & (F_SS | F_MS))) {
            const float gsc = ga * dalpha_dscale(q.sdf, q.alpha, q.sc, A.delta, A.laplace != 0);
            if (flags & F_SS) gs1 += gsc * q.sc / F.scale.x;
            else radius_factor_adjoint(base, q.p, prm, F.scale, gsc * q.sc / q.fac, gpos, gs3, G);
        }
        xform_adjoint(gpos, q.p, x, F, flags, gs1, gs3, gang, gtr);
        if (flags & F_SS) G[2] += gs1;
        else { G[2] += gs3.x; G[3] += gs3.y; G[4] += gs3.z; }
        G[5] += gang.x; G[6] += gang.y; G[7] += gang.z;
        G[8] += gtr.x; G[9] += gtr.y; G[10] += gtr.z;
    }
}

// Adds the slots G of the first nl lanes to dst [S, PG], summed per run of lanes with
// the same stroke key (runs are contiguous), through an LDS transpose: lane l stages
// its slots, then each lane sums one (run, slot) segment and issues one atomic.
template <int NG>
__device__ __forceinline__ void reduce_runs(const RenderArgs& A, const WaveMem& w, int lane, int nl, int key,
                                            const float (&G)[NG], float* dst, int PG) {
#pragma unroll
    for (int g = 0; g < NG; g++) w.scr[lane * (NG + 1) + g] = G[g];
    w.scr[lane * (NG + 1) + NG] = __int_as_float(key);
    const int kprev = __shfl_up(key, 1, 64);
    const bool head = lane < nl && (lane == 0 || key != kprev);
    const unsigned long long hm = __ballot(head);
    const int nseg = __popcll(hm);
    if (head) w.xo[__popcll(hm & ((1ull << lane) - 1ull))] = lane;
    wave_sync();
    const int items = nseg * NG;
    for (int it = lane; it < items; it += 64) {
        const int jseg = it / NG;
        const int g = it - jseg * NG;
        const int lo = w.xo[jseg];
        const int hi = jseg + 1 < nseg ? w.xo[jseg + 1] : nl;
        const int kk = __float_as_int(w.scr[lo * (NG + 1) + NG]);
        float sum = 0.f;
        for (int l = lo; l < hi; l++) sum += w.scr[l * (NG + 1) + g];
        if (sum != 0.f) {
            const int col = slot_column(g, A.table + (int64_t)kk * REC, A.Pmax, A.Q);
            if (col >= 0) atomicAdd(&dst[(int64_t)kk * PG + col], sum);
        }
    }
    wave_sync();
}

// Pending live pairs of the adjoint, queued per wavefront and class in global memory
// so that the gradient groups run 64 pairs of one SDF family at a time, across rays.
// Field-major ring of RCAP slots per class: stroke, ray, sample depth, dL/dalpha,
// dL/d(density param), dL/d(colour) x 3.
constexpr int RCAP = 320;
constexpr int RW = 8;
static_assert(RCAP >= 63 + PBUF, "a window's live pairs of one class fit behind fewer than 64 pending ones");
__host__ __device__ constexpr int64_t ring_floats_per_wave() { return (int64_t)NCLASS * RCAP * RW; }

// global-memory ordering between the lanes of one wavefront (the queue's writes, then
// reads by other lanes)
__device__ __forceinline__ void queue_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

// Evaluates nl (<= 64) queued pairs of one class from slot head on.
template <int CID>
__device__ __forceinline__ void queued_group(const RenderArgs& A, const WaveMem& w, int lane, const float* rq,
                                             int head, int nl, float* dst, int PG) {
    constexpr int NG = 15 + num_color_params(CID);
    float G[NG];
#pragma unroll
    for (int g = 0; g < NG; g++) G[g] = 0.f;
    int k = 0x7fffffff;  // sorts after every stroke
    if (lane < nl) {
        const int slot = (head + lane) % RCAP;
        k = __float_as_int(rq[slot]);
        const int64_t ray = __float_as_int(rq[RCAP + slot]);
        const float t = rq[2 * RCAP + slot];
        const V3 o = ld3(A.origins + 3 * ray), d = ld3(A.dirs + 3 * ray), vd = ld3(A.viewdirs + 3 * ray);
        V3 x = d * t + o;
        if (A.coord_div != 1.f) x = v3(x.x / A.coord_div, x.y / A.coord_div, x.z / A.coord_div);
        const float r = A.cone ? A.radii[ray] * t : A.radii[ray];
        pair_grad<CID>(A, k, x, r, vd, rq[3 * RCAP + slot], rq[4 * RCAP + slot],
                       v3(rq[5 * RCAP + slot], rq[6 * RCAP + slot], rq[7 * RCAP + slot]), G);
    }
    reduce_runs<NG>(A, w, lane, nl, k, G, dst, PG);
}

__host__ __device__ constexpr int pairs_wave_floats(int ng) { return 64 * (ng + 1) + 64; }

// One block: c[0..n) (group counts) replaced by their exclusive prefix, the total in c[n];
// also clears z[0..nz).
__global__ __launch_bounds__(1024) void group_scan_kernel(int* __restrict__ c, int64_t n, float* __restrict__ z,
                                                          int64_t nz) {
    __shared__ int part[1024];
    const int t = threadIdx.x;
    for (int64_t i = t; i < nz; i += 1024) z[i] = 0.f;
    const int64_t per = (n + 1023) / 1024, b0 = min((int64_t)t * per, n), b1 = min(b0 + per, n);
    int sum = 0;
    for (int64_t i = b0; i < b1; i++) sum += c[i];
    part[t] = sum;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {  // inclusive scan of the per-thread sums
        const int v = t >= off ? part[t - off] : 0;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    int run = part[t] - sum;
    for (int64_t i = b0; i < b1; i++) {
        const int v = c[i];
        c[i] = run;
        run += v;
    }
    if (t == 1023) c[n] = part[1023];
}
constexpr int PAIRS_WAVES = 16;  // wavefronts per block of the pair kernel at most (one block per CU)

// Split adjoint, second kernel: the records of render_bwd_kernel<MODE 2>, 64 of one SDF
// family per wavefront, through pair_grad; each stroke's run of lanes is summed into the
// block's rows (LDS, or the global rows) as in the fused kernel.  Does nothing when a
// region overflowed (the fused kernel then takes every ray).
template <int CID>
__global__ __launch_bounds__(64 * PAIRS_WAVES) void render_bwd_pairs_kernel(RenderArgs A, float* __restrict__ partials,
                                                                       int use_lds, float* __restrict__ gdirect) {
    extern __shared__ float smem[];  // [S * PG] accumulators (use_lds) + per-wavefront regions
    constexpr int NG = 15 + num_color_params(CID);
    const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // uniform per wavefront
    const int W = blockDim.x >> 6;
    const int PG = A.Pmax + A.Q + 1;
    const int64_t nacc = use_lds ? (int64_t)A.S * PG : 0;
    float* dst = use_lds ? smem : gdirect;
    for (int64_t i = threadIdx.x; i < nacc; i += blockDim.x) smem[i] = 0.f;
    __syncthreads();
    WaveMem w;
    w.scr = smem + nacc + (int64_t)wave * pairs_wave_floats(NG);
    w.xo = (int*)(w.scr + 64 * (NG + 1));
    // sp_count holds, after grid_scan_kernel, the groups before each region (and the total
    // at index nreg); wavefront q takes groups [q G / Q, (q + 1) G / Q)
    const int64_t nreg = (int64_t)NCLASS * A.sp_nw;
    const int64_t total = *A.sp_ovf == 0 ? A.sp_count[nreg] : 0;
    const int64_t nq = (int64_t)gridDim.x * W, q = (int64_t)blockIdx.x * W + wave;
    const int64_t g0 = total * q / nq, g1 = total * (q + 1) / nq;
    int64_t rg = 0;
    if (g0 < g1) {  // the region of group g0: the last one starting at or before it
        int64_t lo = 0, hi = nreg - 1;
        while (lo < hi) {
            const int64_t mid = (lo + hi + 1) >> 1;
            if (A.sp_count[mid] <= g0) lo = mid;
            else hi = mid - 1;
        }
        rg = lo;
    }
    for (int64_t g = g0; g < g1; g++) {
        while (A.sp_count[rg + 1] <= g) rg++;  // skip regions without groups
        {
            const int64_t gi = g - A.sp_count[rg];
            const float* L = A.sp_list + rg * RW * A.sp_cap + gi * 64 + lane;
            const int k = __float_as_int(L[0]);
            float G[NG];
#pragma unroll
            for (int q = 0; q < NG; q++) G[q] = 0.f;
            if (k >= 0) {
                const int64_t ray = __float_as_int(L[A.sp_cap]);
                const float t = L[2 * A.sp_cap];
                const V3 o = ld3(A.origins + 3 * ray), d = ld3(A.dirs + 3 * ray), vd = ld3(A.viewdirs + 3 * ray);
                V3 x = d * t + o;
                if (A.coord_div != 1.f) x = v3(x.x / A.coord_div, x.y / A.coord_div, x.z / A.coord_div);
                const float r = A.cone ? A.radii[ray] * t : A.radii[ray];
                pair_grad<CID>(A, k, x, r, vd, L[3 * A.sp_cap], L[4 * A.sp_cap],
                               v3(L[5 * A.sp_cap], L[6 * A.sp_cap], L[7 * A.sp_cap]), G);
            }
            // empty records only pad a region's last group, so the valid lanes lead it
            const int nl = __popcll(__ballot(k >= 0));
            reduce_runs<NG>(A, w, lane, nl, k >= 0 ? k : 0x7fffffff, G, dst, PG);
        }
    }
    if (use_lds) {
        __syncthreads();
        float* prow = partials + (int64_t)blockIdx.x * nacc;
        for (int64_t i = threadIdx.x; i < nacc; i += blockDim.x) prow[i] = smem[i];
    }
}

// MODE 0: gradient groups window by window; 1: the pending-pair queues (queues !=
// nullptr), full groups of one family across rays; 2: the first kernel of the split
// adjoint: rays whose hits were all kept and fit the pending list append their live
// pairs' records to this wavefront's region of each family list instead of evaluating
// them (the last group padded with empty records), the other rays are flagged for the
// fused kernel.  With fbflag set, MODE 1 takes only the flagged rays (every ray after
// an overflow).
template <int NS, int CID, int MODE>
__global__ __launch_bounds__(64 * BWAVES) void render_bwd_kernel(RenderArgs A0, const float* __restrict__ g_rgb,
                                                                 const float* __restrict__ g_acc,
                                                                 const float* __restrict__ g_depth,
                                                                 const float* __restrict__ state,
                                                                 float* __restrict__ partials, int use_lds,
                                                                 float* __restrict__ gdirect,
                                                                 float* __restrict__ queues) {
    extern __shared__ float smem[];  // [S * PG] accumulators (use_lds) + per-wavefront regions
    constexpr bool QUEUE = MODE != 0;
    constexpr int NQ = num_color_params(CID);
    constexpr int NG = 15 + NQ;  // gradient slots per pair
    constexpr int NGL = MODE == 2 ? 3 : NG;  // LDS slots per lane (MODE 2: the live-pair codes only)
    constexpr int NPF = 5;       // per-pair values: alpha -> dL/dalpha, colour / dL/d(density, colour)
    const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // uniform per wavefront
    const int W = blockDim.x >> 6;
    const int PG = A0.Pmax + A0.Q + 1;
    // LDS / partial rows are sized for the largest scene; this block's scene may be smaller
    const int64_t nacc_max = use_lds ? (int64_t)A0.S * PG : 0;
    RenderArgs A = A0;
    int64_t ray0 = (int64_t)blockIdx.x * W + wave, ray_end = A0.R, ray_step = (int64_t)gridDim.x * W;
    float* dst = use_lds ? smem : gdirect;
    float* prow = partials ? partials + (int64_t)blockIdx.x * nacc_max : nullptr;
    // this wavefront's pending-pair queues (nullptr: gradient groups run window by window)
    float* queue = MODE == 1 ? queues + ((int64_t)blockIdx.x * W + wave) * ring_floats_per_wave() : nullptr;
    int qhead[NCLASS] = {0, 0, 0}, qtail[NCLASS] = {0, 0, 0};
    if (A0.scenes) {  // a batch: this block works on one scene's rays
        const int b = A0.block_base + (int)blockIdx.x;
        const int sc = A0.scenes[A0.n_scenes * SCENE_REC + b];
        const int* rec = A0.scenes + sc * SCENE_REC;
        scene_strokes(A, rec);
        const int j = b - rec[4];
        ray0 = rec[0] + (int64_t)j * W + wave;
        ray_end = (int64_t)rec[0] + rec[1];
        ray_step = (int64_t)rec[5] * W;
        if (!use_lds) dst = gdirect + (int64_t)rec[2] * PG;
        if (partials) prow = partials + rec[6] + (int64_t)j * rec[3] * PG;
    }
    const int64_t nacc = use_lds ? (int64_t)A.S * PG : 0;
    float* gacc = smem;
    for (int64_t i = threadIdx.x; i < nacc; i += blockDim.x) gacc[i] = 0.f;
    __syncthreads();
    const WaveMem w = wave_mem(smem + nacc_max + wave * wave_mem_floats(NS, NPF, NGL), NS, NPF, NGL);
    const int64_t N = A.R * (int64_t)A.n;
    const bool fb_all = MODE == 1 && A.fbflag && *A.sp_ovf != 0;
    // MODE 2: this wavefront's list regions and how many records each holds
    const int64_t wid = (int64_t)blockIdx.x * W + wave;
    int cur[NCLASS] = {0, 0, 0};
    bool full = false;
    unsigned long long fmask = 0ull;  // MODE 1 with fbflag: flags of this wavefront's next 64 rays
    for (int64_t ray = ray0, it = 0; ray < ray_end; ray += ray_step, it++) {
        if constexpr (MODE == 1) {
            if (A.fbflag && !fb_all) {  // rays done by the split adjoint are skipped
                if ((it & 63) == 0) {
                    const int64_t mine = ray + lane * ray_step;
                    fmask = __ballot(mine < ray_end && A.fbflag[mine] != 0);
                }
                if (!((fmask >> (it & 63)) & 1ull)) continue;
            }
        }
        RayConst rc = load_ray(A, ray);
        float ts[NS];
        bool valid[NS];
        const bool sorted = stage_depths<NS>(A, ray, lane, w, ts, valid);
        if (!sorted) unsorted_range<NS>(A, rc, ts, valid);
        const float* td = A.tdist + ray * (A.n + 1);
        const float dn = sqrtf(rc.d.x * rc.d.x + rc.d.y * rc.d.y + rc.d.z * rc.d.z);
        // ---- ray-march adjoint from the saved per-sample state
        float Dt[NS], Tn[NS], Un[NS], sdl[NS], delta_s[NS];
        V3 Ct[NS], cn[NS];
        bool inr[NS][3];
#pragma unroll
        for (int s = 0; s < NS; s++) {
            const int i = lane + 64 * s;
            if (valid[s]) {
                const int64_t si = ray * A.n + i;
                Dt[s] = state[si];
                Ct[s] = v3(state[N + si], state[2 * N + si], state[3 * N + si]);
                Un[s] = state[4 * N + si];
                Tn[s] = 1.f - Un[s];
                delta_s[s] = (td[i + 1] - td[i]) * dn;
            } else {
                Dt[s] = 0.f; Ct[s] = v3s(0.f); Tn[s] = 1.f; Un[s] = 0.f; delta_s[s] = 0.f;
            }
            const float scl = 1.f / (1e-6f + Un[s]);
            const V3 sv = Ct[s] * scl;
            inr[s][0] = 0.f <= sv.x && sv.x <= 1.f;
            inr[s][1] = 0.f <= sv.y && sv.y <= 1.f;
            inr[s][2] = 0.f <= sv.z && sv.z <= 1.f;
            cn[s] = v3(clampf(sv.x, 0.f, 1.f), clampf(sv.y, 0.f, 1.f), clampf(sv.z, 0.f, 1.f));
            sdl[s] = Dt[s] * delta_s[s];
        }
        float wv[NS], trv[NS];
        float carry = 0.f, acc = 0.f, dsum = 0.f;
#pragma unroll
        for (int s = 0; s < NS; s++) {
            float v = sdl[s];
#pragma unroll
            for (int off = 1; off < 64; off <<= 1) {
                const float u = __shfl_up(v, off, 64);
                if (lane >= off) v += u;
            }
            const float excl = carry + v - sdl[s];
            trv[s] = expf(-excl);
            const float al = (A.opaque && lane + 64 * s == A.n - 1) ? 1.f : 1.f - expf(-sdl[s]);
            wv[s] = valid[s] ? al * trv[s] : 0.f;
            carry += lane_f(v, 63);
            acc += wv[s];
            dsum += wv[s] * ts[s];
        }
        acc = wave_sum(acc);
        dsum = wave_sum(dsum);
        const V3 grgb = ld3(g_rgb + 3 * ray);
        const float gacc_r = g_acc ? g_acc[ray] : 0.f;
        const float gdep = g_depth ? g_depth[ray] : 0.f;
        const V3 bgc = A.bgs ? ld3(A.bgs + 3 * ray) : v3(A.bg0, A.bg1, A.bg2);
        const float bgdot = (1.f - acc >= 0.f) ? dot(grgb, bgc) : 0.f;
        const float accc = fmaxf(acc, 1.1920929e-07f);
        const float dep = dsum / accc;
        const bool dep_in = dep >= td[0] && dep <= td[A.n];
        float gw[NS], gsig[NS];
        float suf_in = 0.f;  // sum over later samples of w_j gw_j
#pragma unroll
        for (int s = 0; s < NS; s++) {
            gw[s] = dot(grgb, cn[s]) + gacc_r - bgdot;
            if (dep_in) gw[s] += gdep * (acc > 1.1920929e-07f ? (ts[s] - dep) / accc : ts[s] / accc);
            if (!valid[s]) gw[s] = 0.f;
        }
        // suffix (exclusive) sums of w*gw over samples after this one
#pragma unroll
        for (int s = NS - 1; s >= 0; s--) {
            float v = wv[s] * gw[s];
#pragma unroll
            for (int off = 1; off < 64; off <<= 1) {
                const float u = __shfl_down(v, off, 64);
                if (lane + off < 64) v += u;
            }
            const float excl = suf_in + v - wv[s] * gw[s];
            // transmittance past the sample (zero past an opaque last interval)
            const float tnext = (A.opaque && lane + 64 * s == A.n - 1) ? 0.f : trv[s] * expf(-sdl[s]);
            gsig[s] = delta_s[s] * (tnext * gw[s] - excl);
            suf_in += lane_f(v, 0);
        }
        // per-sample compositing adjoint inputs
        float gD[NS], O[NS], iOs[NS], Tp[NS], D2[NS];
        V3 gC[NS], C2[NS];
#pragma unroll
        for (int s = 0; s < NS; s++) {
            gD[s] = valid[s] ? gsig[s] : 0.f;
            const V3 gcv = grgb * wv[s];
            gC[s] = v3(inr[s][0] ? gcv.x : 0.f, inr[s][1] ? gcv.y : 0.f, inr[s][2] ? gcv.z : 0.f);
            O[s] = 1e-6f + Un[s];
            iOs[s] = 1.f / O[s];
            Tp[s] = 1.f;
            D2[s] = 0.f;
            C2[s] = v3s(0.f);
        }
        // ---- strokes in compositing order: pending hits are either all kept by the
        //      forward (their pair values are contiguous from gpoff) or all recomputed
        const int64_t rh = ray * (int64_t)A.cap_h;
        const int64_t rp = ray * (int64_t)A.cap_p * (CID == 1 ? 4 : 1);
        int nh = 0, np = 0, gpoff = 0;
        bool gkept = false;
        int d0 = -1, d1 = 0, d2 = 0;  // kept chunk records of chunks (ch & ~63) + lane
        // every chunk kept and the hits fit the pending list: gather them in one pass (a
        // ray's kept hits and pairs are contiguous from its first chunk's offsets)
        bool preloaded = false;
        if (A.hdir && A.nch <= 64) {
            if (lane < A.nch) {
                const int* e = A.hdir + (ray * A.nch_stride + lane) * 3;
                d0 = e[0];
                d1 = e[1];
                d2 = e[2];
            }
            const bool all_kept = __ballot(lane < A.nch && d0 < 0) == 0ull;
            int H = d2;
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) H += __shfl_xor(H, off, 64);
            if (all_kept && H <= HMAX) {
                const int h00 = __builtin_amdgcn_readfirstlane(d0), p00 = __builtin_amdgcn_readfirstlane(d1);
                int last = 0;
                for (int i = lane; i < H; i += 64) {
                    const int* h = A.hbuf + (rh + h00 + i) * 4;
                    w.ck[i] = h[0] & 0x0fffffff;
                    w.ccls[i] = (h[0] >> 28) & 3;
                    w.ci0[i] = h[1];
                    w.ccnt[i] = h[2];
                    w.coff[i] = h[3] - p00;
                    if (i == H - 1) last = h[3] - p00 + h[2];
                }
                wave_sync();
                np = 0;
#pragma unroll
                for (int off = 32; off > 0; off >>= 1) last = max(last, __shfl_xor(last, off, 64));
                np = last;
                nh = H;
                gkept = true;
                gpoff = p00;
                preloaded = true;
            }
        }
        if constexpr (MODE == 2) {
            if (!preloaded) {  // left to the fused kernel
                if (lane == 0) A.fbflag[ray] = 1;
                continue;
            }
        }
        for (int ch = 0; ch < A.nch;) {
            int hoff = -1, kpoff = 0, knc = 0;
            if (preloaded) {
                preloaded = false;
                ch = A.nch;  // the one flush below covers the whole ray
            } else if (A.hdir) {
                if ((ch & 63) == 0) {
                    d0 = -1;
                    if (ch + lane < A.nch) {
                        const int* e = A.hdir + (ray * A.nch_stride + ch + lane) * 3;
                        d0 = e[0];
                        d1 = e[1];
                        d2 = e[2];
                    }
                }
                hoff = __builtin_amdgcn_readlane(d0, ch & 63);
                kpoff = __builtin_amdgcn_readlane(d1, ch & 63);
                knc = __builtin_amdgcn_readlane(d2, ch & 63);
            }
            const bool kept = hoff >= 0;
            bool do_flush = (nh > 0 && kept != gkept) || ch == A.nch;  // flush before switching source
            if (!do_flush) {
                if (nh == 0) {
                    gkept = kept;
                    gpoff = kpoff;
                }
                int nc, tot;
                if (kept) {
                    nc = knc;
                    int end = 0;
                    if (lane < nc) {
                        const int* h = A.hbuf + (rh + hoff + lane) * 4;
                        w.ck[nh + lane] = h[0] & 0x0fffffff;
                        w.ccls[nh + lane] = (h[0] >> 28) & 3;
                        w.ci0[nh + lane] = h[1];
                        w.ccnt[nh + lane] = h[2];
                        w.coff[nh + lane] = np + (h[3] - kpoff);
                        end = h[3] - kpoff + h[2];
                    }
                    tot = nc > 0 ? __shfl(end, nc - 1, 64) : 0;
                    wave_sync();
                } else {
                    if constexpr (MODE != 2) {
                        nc = append_hits<NS>(A, rc, w, sorted, ch * 64, lane, nh, np, tot);
                    } else {
                        nc = 0;
                        tot = 0;
                    }
                }
                nh += nc;
                np += tot;
                ch++;
                do_flush = nh > HMAX - 64 || np >= A.flush_pairs || ch == A.nch;
            }
            if (!do_flush) continue;
            if constexpr (MODE != 2) {
                if (!gkept) nh = refine_pending<NS>(A, rc, w, sorted, nh, lane, np);
            }
            for (int c0 = 0; c0 < nh;) {
                const int wbase = w.coff[c0];
                const int c1 = window_end(w, c0, nh, wbase, lane);
                const int wp = w.coff[c1 - 1] + w.ccnt[c1 - 1] - wbase;
                // (1) alpha (and gradient_rgb colour) per pair: kept values or recomputed
                for (int p0 = 0; gkept && p0 < wp; p0 += 64) {
                    const int p = p0 + lane;
                    if (p < wp) {
                        const float* kp = A.pbuf + rp + gpoff + wbase + p;
                        w.pb[p] = kp[0];
                        if (CID == 1) {
                            w.pb[PBUF + p] = kp[A.cap_p];
                            w.pb[2 * PBUF + p] = kp[2 * (int64_t)A.cap_p];
                            w.pb[3 * PBUF + p] = kp[3 * (int64_t)A.cap_p];
                        }
                    }
                }
                const int hc = c0 + lane;
                const int hcls = hc < c1 ? (w.ccls[hc] & 3) : -1;
                if (MODE != 2 && !gkept) {
                  const int nx = class_sort_hits(w, hc, hcls, lane);
                  const int xt = wp;
                  for (int p0 = 0; p0 < xt; p0 += 64) {
                    const int pp = p0 + lane;
                    if (pp < xt) {
                        const int xr = class_find(w, nx, pp);
                        const int c = w.xh[xr];
                        const int j = pp - w.xo[xr];
                        const int p = w.coff[c] - wbase + j;
                        const int k = w.ck[c];
                        const int i = w.ci0[c] + j;
                        V3 x;
                        float r;
                        sample_point(A, rc, w.ts[i], x, r);
                        const float* rec = A.table + (int64_t)k * REC;
                        const Pair q = eval_pair(rec, ibits(rec[R_BASE]), ibits(rec[R_CURVE]) != 0, x, r, A);
                        w.pb[p] = q.alpha;
                        if (CID == 1) {
                            const V3 col = color_eval_t<1>(q.p, rc.vd, A.cparams + (int64_t)k * A.Q);
                            w.pb[PBUF + p] = col.x;
                            w.pb[2 * PBUF + p] = col.y;
                            w.pb[3 * PBUF + p] = col.z;
                        }
                    }
                  }
                }
                wave_sync();
                // (2) compositing adjoint, strokes front to back; per pair it leaves
                //     dL/dalpha, dL/d(density param) and dL/d(colour) in the staging buffer
                const WindowHits<CID> wh = window_hits<CID>(A, w, c0, c1, wbase, lane, rc.vd);
                for (int c = c0; c < c1; c++) {
                    const int j = c - c0;
                    const int ci = lane_i(wh.ci, j);
                    const int cnn = lane_i(wh.cn, j);
                    const int po = lane_i(wh.po, j);
                    const float dk = lane_f(wh.dk, j);
                    const V3 cu = v3(lane_f(wh.cu.x, j), lane_f(wh.cu.y, j), lane_f(wh.cu.z, j));
#pragma unroll
                    for (int s = 0; s < NS; s++) {
                        const int j = lane + 64 * s - ci;
                        if ((unsigned)j >= (unsigned)cnn) continue;
                        const int q = po + j;
                        const float a = w.pb[q];
                        float ga = 0.f, gdw = 0.f;
                        V3 gcol = v3s(0.f);
                        if (a != 0.f) {
                            const V3 c = (CID == 1) ? v3(w.pb[PBUF + q], w.pb[2 * PBUF + q], w.pb[3 * PBUF + q]) : cu;
                            const float Tk = Tp[s];
                            const float wt = a * Tk;
                            const float Ta = Tk * (1.f - a);
                            D2[s] += dk * wt;
                            C2[s] = C2[s] + c * wt;
                            const float sufD = Dt[s] - D2[s];
                            const V3 sufC = Ct[s] - C2[s];
                            const float inv1 = 1.f / fmaxf(1.f - a, 1e-6f);
                            const float iO = iOs[s];
                            if (A.exact) {
                                ga = gD[s] * (dk * Tk - sufD * inv1);
                                ga += gC[s].x * ((c.x * Tk - sufC.x * inv1) * iO - Ct[s].x * (Tn[s] * inv1) * iO * iO);
                                ga += gC[s].y * ((c.y * Tk - sufC.y * inv1) * iO - Ct[s].y * (Tn[s] * inv1) * iO * iO);
                                ga += gC[s].z * ((c.z * Tk - sufC.z * inv1) * iO - Ct[s].z * (Tn[s] * inv1) * iO * iO);
                            } else {
                                ga = gD[s] * (Ta * dk - sufD) * inv1;
                                const float s2 = inv1 * iO * iO;
                                ga += gC[s].x * (Ta * c.x * (O[s] - a) - sufC.x) * s2;
                                ga += gC[s].y * (Ta * c.y * (O[s] - a) - sufC.y) * s2;
                                ga += gC[s].z * (Ta * c.z * (O[s] - a) - sufC.z) * s2;
                            }
                            Tp[s] = Ta;
                            gdw = gD[s] * wt;
                            gcol = gC[s] * (wt * iO);
                        }
                        w.pb[q] = ga;
                        w.pb[PBUF + q] = gdw;
                        w.pb[2 * PBUF + q] = gcol.x;
                        w.pb[3 * PBUF + q] = gcol.y;
                        w.pb[4 * PBUF + q] = gcol.z;
                    }
                }
                wave_sync();
                // (3) parameter gradients per pair, reduced per stroke.  Only pairs with a
                //     nonzero adjoint contribute (alpha = 0 pairs do not): compact them first,
                //     in order, so that each stroke's pairs stay contiguous
                //     class by class (w.ul: class segments of (hit << 16 | pair), each in pair order)
                int* code = (int*)w.scr;  // [wp] hit << 2 | class of a live pair, -1 otherwise
                int xn[NCLASS] = {0, 0, 0};
                for (int p0 = 0; p0 < wp; p0 += 64) {
                    const int p = p0 + lane;
                    const bool live = p < wp && (w.pb[p] != 0.f || w.pb[PBUF + p] != 0.f ||
                                                 w.pb[2 * PBUF + p] != 0.f || w.pb[3 * PBUF + p] != 0.f ||
                                                 w.pb[4 * PBUF + p] != 0.f);
                    int cd = -1;
                    if (live) {
                        const int c = find_hit(w, c0, c1, wbase, p);
                        cd = (c << 2) | (w.ccls[c] & 3);
                    }
                    if (p < wp) code[p] = cd;
#pragma unroll
                    for (int xc = 0; xc < NCLASS; xc++) xn[xc] += __popcll(__ballot(cd >= 0 && (cd & 3) == xc));
                }
                wave_sync();
                int xs[NCLASS + 1];
                xs[0] = 0;
#pragma unroll
                for (int xc = 0; xc < NCLASS; xc++) xs[xc + 1] = xs[xc] + xn[xc];
                {
                    int xf[NCLASS] = {xs[0], xs[1], xs[2]};
                    for (int p0 = 0; p0 < wp; p0 += 64) {
                        const int p = p0 + lane;
                        const int cd = p < wp ? code[p] : -1;
#pragma unroll
                        for (int xc = 0; xc < NCLASS; xc++) {
                            const bool in = cd >= 0 && (cd & 3) == xc;
                            const unsigned long long m = __ballot(in);
                            if (in) w.ul[xf[xc] + __popcll(m & ((1ull << lane) - 1ull))] = ((cd >> 2) << 16) | p;
                            xf[xc] += __popcll(m);
                        }
                    }
                }
                wave_sync();
                const int nu = xs[NCLASS];
                if constexpr (MODE == 2) {
                    // append the live pairs class by class to this wavefront's regions
#pragma unroll
                    for (int xc = 0; xc < NCLASS; xc++) {
                        const int nx = xs[xc + 1] - xs[xc];
                        if (cur[xc] + nx > A.sp_cap) full = true;
                        if (full) continue;
                        float* L = A.sp_list + ((int64_t)xc * A.sp_nw + wid) * RW * A.sp_cap + cur[xc];
                        for (int u = lane; u < nx; u += 64) {
                            const int e = w.ul[xs[xc] + u];
                            const int p = e & 0xffff, c = e >> 16;
                            L[u] = __int_as_float(w.ck[c]);
                            L[A.sp_cap + u] = __int_as_float((int)ray);
                            L[2 * A.sp_cap + u] = w.ts[w.ci0[c] + (p - (w.coff[c] - wbase))];
                            L[3 * A.sp_cap + u] = w.pb[p];
                            L[4 * A.sp_cap + u] = w.pb[PBUF + p];
                            L[5 * A.sp_cap + u] = w.pb[2 * PBUF + p];
                            L[6 * A.sp_cap + u] = w.pb[3 * PBUF + p];
                            L[7 * A.sp_cap + u] = w.pb[4 * PBUF + p];
                        }
                        cur[xc] += nx;
                    }
                } else if constexpr (QUEUE) {
                    // queue the live pairs class by class behind the pending ones; a class's
                    // full groups of 64 are evaluated as soon as they are there
#pragma unroll
                    for (int xc = 0; xc < NCLASS; xc++) {
                        const int nx = xs[xc + 1] - xs[xc];
                        float* rq = queue + xc * (RCAP * RW);
                        for (int u = lane; u < nx; u += 64) {
                            const int e = w.ul[xs[xc] + u];
                            const int p = e & 0xffff, c = e >> 16;
                            const int slot = (qtail[xc] + u) % RCAP;
                            rq[slot] = __int_as_float(w.ck[c]);
                            rq[RCAP + slot] = __int_as_float((int)ray);
                            rq[2 * RCAP + slot] = w.ts[w.ci0[c] + (p - (w.coff[c] - wbase))];
                            rq[3 * RCAP + slot] = w.pb[p];
                            rq[4 * RCAP + slot] = w.pb[PBUF + p];
                            rq[5 * RCAP + slot] = w.pb[2 * PBUF + p];
                            rq[6 * RCAP + slot] = w.pb[3 * PBUF + p];
                            rq[7 * RCAP + slot] = w.pb[4 * PBUF + p];
                        }
                        qtail[xc] += nx;
                    }
                    queue_sync();
#pragma unroll
                    for (int xc = 0; xc < NCLASS; xc++)
                        for (; qtail[xc] - qhead[xc] >= 64; qhead[xc] += 64)
                            queued_group<CID>(A, w, lane, queue + xc * (RCAP * RW), qhead[xc] % RCAP, 64, dst, PG);
                } else {
                    for (int u0 = 0; u0 < nu; u0 += 64) {
                        const int u = u0 + lane;
                        float G[NG];
    #pragma unroll
                        for (int g = 0; g < NG; g++) G[g] = 0.f;
                        int c = 0x7fffffff;  // sorts after every hit index
                        if (u < nu) {
                            const int p = w.ul[u] & 0xffff;
                            c = w.ul[u] >> 16;
                            const float ga = w.pb[p];
                            const float gdw = w.pb[PBUF + p];
                            const V3 gcol = v3(w.pb[2 * PBUF + p], w.pb[3 * PBUF + p], w.pb[4 * PBUF + p]);
                            {
                                const int k = w.ck[c];
                                const int i = w.ci0[c] + (p - (w.coff[c] - wbase));
                                V3 x;
                                float r;
                                sample_point(A, rc, w.ts[i], x, r);
                                const float* rec = A.table + (int64_t)k * REC;
                                const int base = ibits(rec[R_BASE]);
                                const bool curve = ibits(rec[R_CURVE]) != 0;
                                const uint32_t flags = (uint32_t)ibits(rec[R_FLAGS]);
                                const float* cp = A.cparams + (int64_t)k * A.Q;
                                const float* prm = curve ? rec + R_CP : rec + R_BP;
                                if (curve && A.exact) {
                                    // curves carry no transform: one closest-point solve gives the
                                    // value (hence alpha and dL/dSDF) and the parameter adjoint
                                    const ExactTransfer tr{ga, (A.laplace ? 2.f : 0.5f) / (A.delta * r), A.delta,
                                                           A.laplace != 0};
                                    float sdf, gs;
                                    curve_value_grad(base, x, prm, tr, G, sdf, gs);
                                    color_grad_t<CID>(x, rc.vd, cp, gcol, G + 14, true);
                                    G[14 + NQ] += gdw;
                                } else {
                                const Pair q = eval_pair(rec, base, curve, x, r, A);
                                V3 gpos = color_grad_t<CID>(q.p, rc.vd, cp, gcol, G + 14, A.exact != 0);
                                G[14 + NQ] += gdw;
                                const float gs = ga * dalpha_dsdf(q.sdf, q.alpha, q.sc, A.delta, A.laplace != 0,
                                                                  A.exact != 0);
                                const V3 dsdp = A.exact ? grad_ex(base, q.p, prm, gs, G) : grad_ref(base, q.p, prm, gs, G);
                                gpos = gpos + dsdp * gs;
                                if (!curve) {
                                    float gs1 = 0.f;
                                    V3 gs3 = v3s(0.f), gang = v3s(0.f), gtr = v3s(0.f);
                                    Frame F;
                                    F.trans = ld3(rec + R_TR);
                                    F.angle = ld3(rec + R_ANG);
                                    F.sz = rec[R_SZ];
                                    F.cz = rec[R_CZ];
                                    F.scale = ld3(rec + R_SC);
    #pragma unroll
                                    for (int ii = 0; ii < 3; ii++)
    #pragma unroll
                                        for (int jj = 0; jj < 3; jj++) F.R.r[ii][jj] = rec[R_ROT + 3 * ii + jj];
                                    if (A.exact && A.isr && (flags & (F_SS | F_MS))) {
                                        const float gsc = ga * dalpha_dscale(q.sdf, q.alpha, q.sc, A.delta, A.laplace != 0);
                                        if (flags & F_SS) gs1 += gsc * q.sc / F.scale.x;
                                        else radius_factor_adjoint(base, q.p, prm, F.scale, gsc * q.sc / q.fac, gpos, gs3, G);
                                    }
                                    xform_adjoint(gpos, q.p, x, F, flags, gs1, gs3, gang, gtr);
                                    if (flags & F_SS) G[2] += gs1;
                                    else { G[2] += gs3.x; G[3] += gs3.y; G[4] += gs3.z; }
                                    G[5] += gang.x; G[6] += gang.y; G[7] += gang.z;
                                    G[8] += gtr.x; G[9] += gtr.y; G[10] += gtr.z;
                                }
                                }
                            }
                        }
                        // segmented sum over the lanes of each hit through an LDS transpose (the
                        // lanes of one hit are contiguous; w.xo lists the first lane of each run)
                        const int nl = min(64, nu - u0);
    #pragma unroll
                        for (int g = 0; g < NG; g++) w.scr[lane * (NG + 1) + g] = G[g];
                        w.scr[lane * (NG + 1) + NG] = __int_as_float(c);
                        const int cprev = __shfl_up(c, 1, 64);
                        const bool head = lane < nl && (lane == 0 || c != cprev);
                        const unsigned long long hm = __ballot(head);
                        const int nseg = __popcll(hm);
                        if (head) w.xo[__popcll(hm & ((1ull << lane) - 1ull))] = lane;
                        wave_sync();
                        const int items = nseg * NG;
                        for (int it = lane; it < items; it += 64) {
                            const int jseg = it / NG;
                            const int g = it - jseg * NG;
                            const int lo = w.xo[jseg];
                            const int hi = jseg + 1 < nseg ? w.xo[jseg + 1] : nl;
                            const int cc = __float_as_int(w.scr[lo * (NG + 1) + NG]);
                            float sum = 0.f;
                            for (int l = lo; l < hi; l++) sum += w.scr[l * (NG + 1) + g];
                            if (sum != 0.f) {
                                const int k = w.ck[cc];
                                const int col = slot_column(g, A.table + (int64_t)k * REC, A.Pmax, A.Q);
                                if (col >= 0) atomicAdd(&dst[(int64_t)k * PG + col], sum);
                            }
                        }
                        wave_sync();
                    }
                }
                c0 = c1;
            }
            nh = np = 0;
        }
    }
    if constexpr (MODE == 2) {  // pad each region to whole groups; a full region sends every ray to MODE 1
#pragma unroll
        for (int xc = 0; xc < NCLASS; xc++) {
            const int padded = (cur[xc] + 63) / 64 * 64;  // <= sp_cap, a multiple of 64
            float* L = A.sp_list + ((int64_t)xc * A.sp_nw + wid) * RW * A.sp_cap;
            if (!full && cur[xc] + lane < padded) L[cur[xc] + lane] = __int_as_float(-1);
            if (lane == 0) A.sp_count[(int64_t)xc * A.sp_nw + wid] = full ? 0 : padded / 64;
        }
        if (full && lane == 0) *A.sp_ovf = 1;
    } else if constexpr (QUEUE) {  // the partial groups left in the queues
#pragma unroll
        for (int xc = 0; xc < NCLASS; xc++)
            for (; qtail[xc] > qhead[xc]; qhead[xc] += 64)
                queued_group<CID>(A, w, lane, queue + xc * (RCAP * RW), qhead[xc] % RCAP,
                                  min(64, qtail[xc] - qhead[xc]), dst, PG);
    }
    if (use_lds) {
        __syncthreads();
        for (int64_t i = threadIdx.x; i < nacc; i += blockDim.x) prow[i] = gacc[i];
    }
}

// Sum per-block partial rows into the parameter-gradient tensors (+=).  blockIdx.y
// takes a slice of RED_PARTS consecutive rows; slices meet in float atomics (one
// slice: a plain read-modify-write).
constexpr int RED_PARTS = 16;

__device__ __forceinline__ void add_grad(float* p, float v, bool atomic) {
    if (atomic) atomicAdd(p, v);
    else *p += v;
}

__global__ void render_reduce_kernel(const float* __restrict__ partials, int nparts, int S, int Pmax, int Q,
                                     float* __restrict__ gshape, float* __restrict__ gcolor,
                                     float* __restrict__ gdens) {
    const int PG = Pmax + Q + 1;
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)S * PG) return;
    const int b0 = blockIdx.y * RED_PARTS, b1 = min(nparts, b0 + RED_PARTS);
    float s = 0.f;
#pragma unroll 8
    for (int b = b0; b < b1; b++) s += partials[(int64_t)b * S * PG + i];
    const bool atomic = gridDim.y > 1;
    int k = (int)(i / PG), c = (int)(i - (int64_t)k * PG);
    if (c < Pmax) add_grad(&gshape[(int64_t)k * Pmax + c], s, atomic);
    else if (c < Pmax + Q) add_grad(&gcolor[(int64_t)k * Q + (c - Pmax)], s, atomic);
    else add_grad(&gdens[k], s, atomic);
}

// batch of scenes: stroke k of scene s sums the partial rows of s's blocks
// (rows of s's S * PG floats from the scene's offset), or, for a scene whose
// accumulators did not fit in LDS, reads its row of the global accumulators
__global__ void render_batch_reduce_kernel(const float* __restrict__ partials, const float* __restrict__ direct,
                                           const int* __restrict__ scenes, int n_scenes, int S_total, int Pmax,
                                           int Q, float* __restrict__ gshape, float* __restrict__ gcolor,
                                           float* __restrict__ gdens) {
    const int PG = Pmax + Q + 1;
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)S_total * PG) return;
    const int k = (int)(i / PG), c = (int)(i - (int64_t)k * PG);
    const int* rec = scenes + find_scene(scenes, n_scenes, 2, k) * SCENE_REC;
    // blockIdx.y takes a slice of RED_PARTS of the scene's blocks
    const int b0 = blockIdx.y * RED_PARTS, b1 = min(rec[7] == 0 ? 1 : rec[5], b0 + RED_PARTS);
    if (b0 >= b1) return;
    float s = 0.f;
    if (rec[7] == 0) {
        if (rec[5] > 0) s = direct[i];
    } else {
        const int64_t row = (int64_t)rec[3] * PG;
        const float* p = partials + rec[6] + (int64_t)(k - rec[2]) * PG + c;
        for (int b = b0; b < b1; b++) s += p[b * row];
    }
    const bool atomic = gridDim.y > 1;
    if (c < Pmax) add_grad(&gshape[(int64_t)k * Pmax + c], s, atomic);
    else if (c < Pmax + Q) add_grad(&gcolor[(int64_t)k * Q + (c - Pmax)], s, atomic);
    else add_grad(&gdens[k], s, atomic);
}

__global__ void fill_zero_kernel(float* __restrict__ p, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.f;
}

static int g_num_cu = 0;
static int g_max_lds = 0;

// Launch-shape parameters chosen by the autotuner (n3ds_render_tuning): pending pairs
// that trigger an evaluation window, backward persistent blocks per resident block,
// the batch backward's blocks per resident block, and chunks per culling task.
// TUNE_BWD_QUEUE: the adjoint queues live pairs per SDF family across rays (1) or
// evaluates each window's pairs as they come (0).  TUNE_CULL_GRID: the culling pass of
// one scene tests only the strokes of its ray's grid cells (1) or every stroke (0).
// TUNE_BWD_SPLIT: with kept hits, the adjoint runs as a per-ray kernel that lists the
// live pairs' records and a kernel over the records (1), or as one kernel (0);
// TUNE_SPLIT_CAP: list records per ray and SDF family (an overflow falls back to the one
// kernel for every ray).
enum {
    TUNE_FLUSH_PAIRS = 0, TUNE_BWD_OVERSUB = 1, TUNE_BATCH_OVERSUB = 2, TUNE_CULL_GROUP = 3, TUNE_BWD_QUEUE = 4,
    TUNE_CULL_GRID = 5, TUNE_BWD_SPLIT = 6, TUNE_SPLIT_CAP = 7, TUNE_COUNT = 8
};
static int g_tune[TUNE_COUNT] = {FLUSH_PAIRS, 1, 8, 1, 1, 1, 1, 32};
static const int g_tune_lo[TUNE_COUNT] = {16, 1, 1, 1, 0, 0, 0, 1};
static const int g_tune_hi[TUNE_COUNT] = {PBUF, 8, 64, 16, 1, 1, 1, 1024};

typedef void (*FwdKernel)(RenderArgs, float*, float*, float*, float*, unsigned long long*);
typedef void (*BwdKernel)(RenderArgs, const float*, const float*, const float*, const float*, float*, int, float*,
                          float*);

static FwdKernel fwd_pre_kernel_ptr(int ns, int cid) {
    static const FwdKernel tab[4][4] = {
        {render_fwd_pre_kernel<1, 0>, render_fwd_pre_kernel<1, 1>, render_fwd_pre_kernel<1, 2>,
         render_fwd_pre_kernel<1, 3>},
        {render_fwd_pre_kernel<2, 0>, render_fwd_pre_kernel<2, 1>, render_fwd_pre_kernel<2, 2>,
         render_fwd_pre_kernel<2, 3>},
        {render_fwd_pre_kernel<3, 0>, render_fwd_pre_kernel<3, 1>, render_fwd_pre_kernel<3, 2>,
         render_fwd_pre_kernel<3, 3>},
        {render_fwd_pre_kernel<4, 0>, render_fwd_pre_kernel<4, 1>, render_fwd_pre_kernel<4, 2>,
         render_fwd_pre_kernel<4, 3>}};
    return tab[ns - 1][cid];
}

static FwdKernel fwd_kernel_ptr(int ns, int cid) {
    static const FwdKernel tab[4][4] = {
        {render_fwd_kernel<1, 0>, render_fwd_kernel<1, 1>, render_fwd_kernel<1, 2>, render_fwd_kernel<1, 3>},
        {render_fwd_kernel<2, 0>, render_fwd_kernel<2, 1>, render_fwd_kernel<2, 2>, render_fwd_kernel<2, 3>},
        {render_fwd_kernel<3, 0>, render_fwd_kernel<3, 1>, render_fwd_kernel<3, 2>, render_fwd_kernel<3, 3>},
        {render_fwd_kernel<4, 0>, render_fwd_kernel<4, 1>, render_fwd_kernel<4, 2>, render_fwd_kernel<4, 3>}};
    return tab[ns - 1][cid];
}

#define N3DS_BWD_ROW(NS, M) \
    {render_bwd_kernel<NS, 0, M>, render_bwd_kernel<NS, 1, M>, render_bwd_kernel<NS, 2, M>, render_bwd_kernel<NS, 3, M>}
// mode 0: window by window, 1: queued groups, 2: the split adjoint's first kernel
static BwdKernel bwd_kernel_ptr(int ns, int cid, int mode = 0) {
    static const BwdKernel tab[3][4][4] = {
        {N3DS_BWD_ROW(1, 0), N3DS_BWD_ROW(2, 0), N3DS_BWD_ROW(3, 0), N3DS_BWD_ROW(4, 0)},
        {N3DS_BWD_ROW(1, 1), N3DS_BWD_ROW(2, 1), N3DS_BWD_ROW(3, 1), N3DS_BWD_ROW(4, 1)},
        {N3DS_BWD_ROW(1, 2), N3DS_BWD_ROW(2, 2), N3DS_BWD_ROW(3, 2), N3DS_BWD_ROW(4, 2)}};
    return tab[mode][ns - 1][cid];
}

static int fwd_lds_bytes(int ns, int cid) { return FWAVES * wave_mem_floats(ns, cid == 1 ? 4 : 1, 0) * 4; }

static int bwd_wave_floats(int ns, int cid) { return wave_mem_floats(ns, 5, 15 + num_color_params(cid)); }

// upper bound on persistent backward blocks: 8 per CU, never more than one per ray
static int64_t max_persistent_blocks(int64_t n_rays) {
    int64_t nb = (int64_t)g_num_cu * 8;
    if (nb > n_rays) nb = n_rays;
    return nb < 1 ? 1 : nb;
}

// Backward launch shape: wavefronts per block and whether the [S, PG] accumulators fit in LDS.
struct BwdShape {
    int waves, use_lds;
    int64_t lds_bytes;
};

static BwdShape bwd_shape(int ns, int cid, int S, int PG) {
    BwdShape b;
    const int64_t wave_bytes = (int64_t)bwd_wave_floats(ns, cid) * 4;
    const int64_t acc_bytes = (int64_t)S * PG * 4;
    for (int wv = BWAVES; wv >= 2; wv >>= 1) {
        if (acc_bytes + wv * wave_bytes <= g_max_lds) {
            b.waves = wv;
            b.use_lds = 1;
            b.lds_bytes = acc_bytes + wv * wave_bytes;
            return b;
        }
    }
    b.waves = BWAVES;  // accumulators in global memory
    b.use_lds = 0;
    b.lds_bytes = BWAVES * wave_bytes;
    return b;
}

static int device_info() {
    if (g_num_cu == 0) {
        int dev;
        N3DS_HIP_CHECK(hipGetDevice(&dev));
        N3DS_HIP_CHECK(hipDeviceGetAttribute(&g_num_cu, hipDeviceAttributeMultiprocessorCount, dev));
        N3DS_HIP_CHECK(hipDeviceGetAttribute(&g_max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
    }
    return N3DS_OK;
}

typedef void (*BwdKernelT)(RenderArgs, const float*, const float*, const float*, const float*, float*, int, float*,
                           float*);

// Persistent backward blocks for n_rays rays: as many as fit on the device at once
// (registers and LDS) times the oversubscription setting, at most one per bs.waves rays.
static int bwd_blocks(BwdKernelT kern, const BwdShape& bs, int64_t n_rays, int64_t* nb_out) {
    if (bs.lds_bytes > 65536)
        N3DS_HIP_CHECK(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize,
                                           (int)bs.lds_bytes));
    int64_t nb = max_persistent_blocks(n_rays);
    const int64_t need = (n_rays + bs.waves - 1) / bs.waves;
    if (nb > need) nb = need;
    int per_cu = 0;
    N3DS_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, 64 * bs.waves, (size_t)bs.lds_bytes));
    if (per_cu < 1) per_cu = 1;
    const int64_t fit = (int64_t)g_num_cu * per_cu * g_tune[TUNE_BWD_OVERSUB];
    if (fit < nb) nb = fit;
    *nb_out = nb < 1 ? 1 : nb;
    return N3DS_OK;
}

}  // namespace n3ds

using namespace n3ds;

static int render_check(int64_t R, int n, int S, int Pmax, int Q, int cid) {
    N3DS_REQUIRE(R >= 0 && S >= 0 && S < (1 << 28), "bad ray / stroke count");
    N3DS_REQUIRE(n >= 2 && n <= 256, "samples per ray must be in [2, 256] (got %d)", n);
    N3DS_REQUIRE(Pmax >= 1 && Pmax <= 23, "shape_params width %d out of range", Pmax);
    N3DS_REQUIRE(cid >= 0 && cid < 4, "color_id %d out of range", cid);
    N3DS_REQUIRE(Q == num_color_params(cid), "color_params width %d does not match color_id %d", Q, cid);
    return N3DS_OK;
}

static RenderArgs make_args(const float* origins, const float* dirs, const float* viewdirs, const float* radii,
                            const float* tdist, const float* table, const float* cparams, const float* dparams,
                            int64_t R, int n, int S, int Pmax, int Q, int cid, float delta, int laplace, int isr,
                            int cone, float coord_div, const float* bg, const float* bg_rgbs, int opaque,
                            float cull_eps, int exact) {
    RenderArgs A;
    A.origins = origins; A.dirs = dirs; A.viewdirs = viewdirs; A.radii = radii; A.tdist = tdist; A.table = table;
    A.cparams = cparams; A.dparams = dparams; A.R = R; A.n = n; A.S = S; A.Q = Q; A.Pmax = Pmax; A.cid = cid;
    A.delta = delta; A.laplace = laplace; A.isr = isr; A.cone = cone; A.exact = exact;
    A.coord_div = coord_div; A.bg0 = bg[0]; A.bg1 = bg[1]; A.bg2 = bg[2];
    A.bgs = bg_rgbs; A.opaque = opaque;
    if (laplace) {
        float eps = cull_eps > 0.f ? cull_eps : 1e-7f;
        A.tail = fmaxf(1.f, 0.5f * logf(0.5f / eps)) * 1.001f;
    } else {
        A.tail = 1.001f;
    }
    A.hdir = nullptr;
    A.hbuf = nullptr;
    A.pbuf = nullptr;
    A.cap_h = A.cap_p = 0;
    A.nch = (S + 63) / 64;
    A.nch_stride = A.nch;
    A.scenes = nullptr;
    A.n_scenes = 0;
    A.block_base = 0;
    A.ccounts = nullptr;
    A.cslots = nullptr;
    A.flush_pairs = g_tune[TUNE_FLUSH_PAIRS];
    A.cull_group = g_tune[TUNE_CULL_GROUP];
    A.ntask = 0;
    A.sp_list = nullptr;
    A.sp_count = nullptr;
    A.sp_ovf = nullptr;
    A.fbflag = nullptr;
    A.sp_cap = 0;
    A.sp_nw = 0;
    return A;
}

static int set_keep(RenderArgs& A, int* hit_dir, int* hit_buf, float* pair_buf, int cap_hits, int cap_pairs) {
    if (!hit_dir) return N3DS_OK;
    N3DS_REQUIRE(hit_buf && pair_buf && cap_hits > 0 && cap_pairs > 0,
                 "kept hits need hit_buf, pair_buf and positive capacities");
    A.hdir = hit_dir;
    A.hbuf = hit_buf;
    A.pbuf = pair_buf;
    A.cap_h = cap_hits;
    A.cap_p = cap_pairs;
    return N3DS_OK;
}

N3DS_EXPORT int64_t n3ds_render_table_floats(int n_strokes) { return (int64_t)n_strokes * REC; }

N3DS_EXPORT int n3ds_render_tuning(int param, int value) {
    if (param < 0 || param >= TUNE_COUNT) {
        set_error("unknown tuning parameter %d", param);
        return -1;
    }
    const int old = g_tune[param];
    if (value >= 0) {
        if (value < g_tune_lo[param] || value > g_tune_hi[param]) {
            set_error("tuning parameter %d: %d outside [%d, %d]", param, value, g_tune_lo[param], g_tune_hi[param]);
            return -1;
        }
        g_tune[param] = value;
    }
    return old;
}

typedef void (*CullKernel)(RenderArgs, int*, int*);

static CullKernel cull_kernel_ptr(int ns) {
    static const CullKernel tab[4] = {render_cull_kernel<1>, render_cull_kernel<2>, render_cull_kernel<3>,
                                      render_cull_kernel<4>};
    return tab[ns - 1];
}

static int64_t cull_workspace_bytes(int64_t n_rays, int nch) { return n_rays * nch * 65 * 4; }

// Runs the culling pass into the caller's workspace (if given) and points A at its hits.
static int launch_cull(RenderArgs& A, int ns, void* workspace, int64_t workspace_bytes, hipStream_t s,
                       int64_t ntask) {
    if (!workspace) return N3DS_OK;
    A.ntask = ntask;
    N3DS_REQUIRE(workspace_bytes >= cull_workspace_bytes(A.R, A.nch_stride),
                 "forward workspace has %lld bytes, need %lld", (long long)workspace_bytes,
                 (long long)cull_workspace_bytes(A.R, A.nch_stride));
    N3DS_REQUIRE(ntask / CWAVES < (1ll << 31), "too many (ray, chunk) culling tasks");
    if (ntask == 0) return N3DS_OK;
    int* counts = (int*)workspace;
    int* slots = counts + A.R * A.nch_stride;
    hipLaunchKernelGGL(cull_kernel_ptr(ns), dim3((unsigned)((ntask + CWAVES - 1) / CWAVES)), dim3(64 * CWAVES),
                       CWAVES * cull_wave_floats(ns) * 4, s, A, counts, slots);
    N3DS_LAUNCH_CHECK();
    A.ccounts = counts;
    A.cslots = slots;
    return N3DS_OK;
}

// Grid of the grid-assisted culling pass, behind the culling workspace (256-byte
// aligned pieces): parameter block, per-block margins, per-ray segments, cell masks
// [res^3, nch] and per-ray masks [R, nch].
struct GridWs {
    float* gp;
    float* part;
    float* seg;
    unsigned long long* cells;
    unsigned long long* rmask;
    int res;
    int64_t nparts, ncell_words, bytes;
};

static int64_t align256(int64_t b) { return (b + 255) & ~(int64_t)255; }

// byte offsets of the pieces (parameters, margins, segments, cells, ray masks, end)
static void grid_offsets(int64_t n_rays, int nch, int64_t* off) {
    const int res = grid_res(nch);
    off[0] = 0;
    off[1] = off[0] + align256(GP_COUNT * 4);
    off[2] = off[1] + align256((n_rays + GRID_RAYS_PER_BLOCK - 1) / GRID_RAYS_PER_BLOCK * 4);
    off[3] = off[2] + align256(n_rays * 12);
    off[4] = off[3] + align256((int64_t)res * res * res * nch * 8);
    off[5] = off[4] + align256(n_rays * nch * 8);
}

static GridWs grid_layout(char* base, int64_t n_rays, int nch) {
    GridWs g;
    int64_t off[6];
    grid_offsets(n_rays, nch, off);
    g.res = grid_res(nch);
    g.nparts = (n_rays + GRID_RAYS_PER_BLOCK - 1) / GRID_RAYS_PER_BLOCK;
    g.ncell_words = (int64_t)g.res * g.res * g.res * nch;
    g.gp = (float*)(base + off[0]);
    g.part = (float*)(base + off[1]);
    g.seg = (float*)(base + off[2]);
    g.cells = (unsigned long long*)(base + off[3]);
    g.rmask = (unsigned long long*)(base + off[4]);
    g.bytes = off[5];
    return g;
}

static int64_t grid_workspace_bytes(int64_t n_rays, int nch) {
    if (nch < 1 || nch > GRID_MAX_NCH) return 0;
    int64_t off[6];
    grid_offsets(n_rays, nch, off);
    return off[5];
}

typedef void (*CullGridKernel)(RenderArgs, const unsigned long long*, int*, int*);

// The culling pass of one scene with the grid: rays, setup, insertion, walk, then the
// candidate culling.  Launches nothing (and leaves *used false) when the grid does not
// apply: a batch, no strokes, more than GRID_MAX_NCH chunks, a workspace without room.
static int launch_cull_grid(RenderArgs& A, int ns, void* workspace, int64_t workspace_bytes, hipStream_t s,
                            bool* used) {
    *used = false;
    const int64_t cb = align256(cull_workspace_bytes(A.R, A.nch_stride));
    if (!workspace || !g_tune[TUNE_CULL_GRID] || A.scenes || A.nch < 1 || A.nch > GRID_MAX_NCH ||
        A.nch_stride != A.nch || workspace_bytes < cb + grid_workspace_bytes(A.R, A.nch))
        return N3DS_OK;
    N3DS_REQUIRE(A.R / CWAVES < (1ll << 31) && A.R * (int64_t)(A.nch + 7) / 8 * 64 / 256 < (1ll << 31) &&
                 (int64_t)A.S * NPIECE / 4 < (1ll << 31), "too many rays or strokes for the culling grid");
    int* counts = (int*)workspace;
    int* slots = counts + A.R * A.nch_stride;
    const GridWs g = grid_layout((char*)workspace + cb, A.R, A.nch);
    static const CullGridKernel ck[4] = {render_cull_grid_kernel<1>, render_cull_grid_kernel<2>,
                                         render_cull_grid_kernel<3>, render_cull_grid_kernel<4>};
    hipLaunchKernelGGL(grid_rays_kernel, dim3((unsigned)g.nparts), dim3(GRID_RAYS_PER_BLOCK), 0, s, A, g.seg, g.part,
                       g.cells, g.ncell_words);
    N3DS_LAUNCH_CHECK();
    hipLaunchKernelGGL(grid_setup_kernel, dim3(1), dim3(GRID_SETUP_THREADS), 0, s, A, (const float*)g.part,
                       (int)g.nparts, g.res, g.gp);
    N3DS_LAUNCH_CHECK();
    const int64_t items = (int64_t)A.S * NPIECE;
    hipLaunchKernelGGL(grid_insert_kernel, dim3((unsigned)((items + 3) / 4)), dim3(256), 0, s, A,
                       (const float*)g.gp, g.cells);
    N3DS_LAUNCH_CHECK();
    const int64_t walks = (A.R + GRID_WALK_RAYS - 1) / GRID_WALK_RAYS * ((A.nch + GRID_WALK_CH - 1) / GRID_WALK_CH);
    hipLaunchKernelGGL(grid_walk_kernel, dim3((unsigned)((walks + GRID_WALK_WAVES - 1) / GRID_WALK_WAVES)),
                       dim3(64 * GRID_WALK_WAVES), 0, s, A,
                       (const float*)g.gp, (const float*)g.seg, (const unsigned long long*)g.cells, g.rmask);
    N3DS_LAUNCH_CHECK();
    hipLaunchKernelGGL(ck[ns - 1], dim3((unsigned)((A.R + CWAVES - 1) / CWAVES)), dim3(64 * CWAVES),
                       CWAVES * cullg_wave_floats(ns) * 4, s, A, (const unsigned long long*)g.rmask, counts, slots);
    N3DS_LAUNCH_CHECK();
    A.ccounts = counts;
    A.cslots = slots;
    *used = true;
    return N3DS_OK;
}

N3DS_EXPORT int n3ds_render_extras(const float* state, const float* tdist, const float* directions,
                                   const float* t_far, int64_t n_rays, int n_samples, int opaque_background,
                                   float* extras_out, void* stream) {
    N3DS_REQUIRE(n_rays >= 0 && n_samples >= 1, "bad ray sizes");
    N3DS_REQUIRE(n_rays == 0 || (state && tdist && directions && t_far && extras_out), "a required pointer is NULL");
    if (n_rays == 0) return N3DS_OK;
    hipLaunchKernelGGL(render_extras_kernel, dim3((unsigned)((n_rays + XWAVES - 1) / XWAVES)), dim3(64 * XWAVES), 0,
                       (hipStream_t)stream, state, tdist, directions, t_far, n_rays, n_samples, opaque_background,
                       extras_out);
    N3DS_LAUNCH_CHECK();
    return N3DS_OK;
}

N3DS_EXPORT int64_t n3ds_render_forward_workspace_bytes(int64_t n_rays, int max_strokes) {
    if (n_rays < 0 || max_strokes < 0) return -1;
    const int nch = (max_strokes + 63) / 64;
    return align256(cull_workspace_bytes(n_rays, nch)) + grid_workspace_bytes(n_rays, nch);
}

N3DS_EXPORT int n3ds_render_prepare(float* table, const int32_t* stroke_ids, const float* shape_params,
                                    int n_strokes, int shape_width, int* status, void* stream) {
    N3DS_REQUIRE(n_strokes >= 0 && shape_width >= 1 && shape_width <= 23, "bad stroke table sizes");
    if (n_strokes == 0) return N3DS_OK;
    static_assert(64 % NPIECE == 0 && 128 % NPIECE == 0, "a stroke's lanes share a wavefront");
    hipLaunchKernelGGL(render_prep_kernel, dim3((unsigned)(((int64_t)n_strokes * NPIECE + 127) / 128)), dim3(128), 0,
                       (hipStream_t)stream, table,
                       stroke_ids, shape_params, n_strokes, shape_width, status);
    N3DS_LAUNCH_CHECK();
    return N3DS_OK;
}

N3DS_EXPORT int n3ds_render_forward(const float* origins, const float* directions, const float* viewdirs,
                                    const float* radii, const float* tdist, const float* table,
                                    const float* color_params, const float* density_params, int64_t n_rays,
                                    int n_samples, int n_strokes, int shape_width, int color_width, int color_id,
                                    float sdf_delta, int use_laplace_transform, int inv_scale_radius, int cone,
                                    float coord_div, const float* bg_color, const float* bg_rgbs, int opaque_background, float cull_eps, float* rgb_out,
                                    float* acc_out, float* depth_out, float* state_out,
                                    unsigned long long* counters, int* hit_dir, int* hit_buf, float* pair_buf,
                                    int cap_hits, int cap_pairs, void* workspace, int64_t workspace_bytes,
                                    void* stream) {
    int st = render_check(n_rays, n_samples, n_strokes, shape_width, color_width, color_id);
    if (st) return st;
    if (n_rays == 0) return N3DS_OK;
    RenderArgs A = make_args(origins, directions, viewdirs, radii, tdist, table, color_params, density_params,
                             n_rays, n_samples, n_strokes, shape_width, color_width, color_id, sdf_delta,
                             use_laplace_transform, inv_scale_radius, cone, coord_div, bg_color, bg_rgbs, opaque_background, cull_eps, 0);
    st = set_keep(A, hit_dir, hit_buf, pair_buf, cap_hits, cap_pairs);
    if (st) return st;
    dim3 grid((unsigned)((n_rays + FWAVES - 1) / FWAVES)), block(64 * FWAVES);
    const int ns = (n_samples + 63) / 64;
    bool gridded = false;
    st = launch_cull_grid(A, ns, workspace, workspace_bytes, (hipStream_t)stream, &gridded);
    if (st) return st;
    if (!gridded) {
        st = launch_cull(A, ns, workspace, workspace_bytes, (hipStream_t)stream,
                         n_rays * (int64_t)((A.nch_stride + A.cull_group - 1) / A.cull_group));
        if (st) return st;
    }
    hipLaunchKernelGGL(A.ccounts ? fwd_pre_kernel_ptr(ns, color_id) : fwd_kernel_ptr(ns, color_id), grid, block,
                       fwd_lds_bytes(ns, color_id), (hipStream_t)stream,
                       A, rgb_out, acc_out, depth_out, state_out, counters);
    N3DS_LAUNCH_CHECK();
    return N3DS_OK;
}

// Split adjoint launch plan: the per-ray kernel (MODE 2, SPLIT_WAVES wavefronts per
// block, no gradient rows), the pair kernel and the fused kernel for flagged rays (both
// with per-block gradient rows in LDS), and the workspace pieces in floats: gradient
// rows of both, the two kernels' queues, the family lists, then the counters and flags.
constexpr int SPLIT_WAVES = 4;

struct SplitPlan {
    bool ok;
    BwdKernel k1, kf;
    int64_t nb1, nb2, nbf;
    int64_t lds1, lds2;
    int pw;  // wavefronts per block of the pair kernel
    BwdShape bsf;
    int64_t cap;  // records per region (a multiple of 64)
    int64_t off_rows, off_qf, off_lists, off_flags, total;
};

typedef void (*PairsKernel)(RenderArgs, float*, int, float*);

static PairsKernel pairs_kernel_ptr(int cid) {
    static const PairsKernel tab[4] = {render_bwd_pairs_kernel<0>, render_bwd_pairs_kernel<1>,
                                       render_bwd_pairs_kernel<2>, render_bwd_pairs_kernel<3>};
    return tab[cid];
}

static int split_plan(int ns, int cid, int64_t n_rays, int S, int PG, SplitPlan* sp) {
    sp->ok = false;
    if (!g_tune[TUNE_BWD_SPLIT] || !g_tune[TUNE_BWD_QUEUE]) return N3DS_OK;
    // the one-kernel adjoint for the flagged rays adds into one global row set
    sp->bsf.waves = BWAVES;
    sp->bsf.use_lds = 0;
    sp->bsf.lds_bytes = (int64_t)BWAVES * bwd_wave_floats(ns, cid) * 4;
    if (sp->bsf.lds_bytes > g_max_lds) return N3DS_OK;
    const int ng = 15 + num_color_params(cid);
    // the pair kernel: as many wavefronts per block as the rows leave LDS for
    sp->pw = 0;
    for (int wv = PAIRS_WAVES; wv >= 4 && !sp->pw; wv -= 4)
        if ((int64_t)S * PG * 4 + (int64_t)wv * pairs_wave_floats(ng) * 4 <= g_max_lds) sp->pw = wv;
    if (!sp->pw) return N3DS_OK;
    sp->lds2 = (int64_t)S * PG * 4 + (int64_t)sp->pw * pairs_wave_floats(ng) * 4;
    sp->lds1 = (int64_t)SPLIT_WAVES * wave_mem_floats(ns, 5, 3) * 4;
    sp->k1 = bwd_kernel_ptr(ns, cid, 2);
    sp->kf = bwd_kernel_ptr(ns, cid, 1);
    if (sp->lds1 > 65536)
        N3DS_HIP_CHECK(hipFuncSetAttribute((const void*)sp->k1, hipFuncAttributeMaxDynamicSharedMemorySize,
                                           (int)sp->lds1));
    if (sp->lds2 > 65536)
        N3DS_HIP_CHECK(hipFuncSetAttribute((const void*)pairs_kernel_ptr(cid),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)sp->lds2));
    int per_cu = 0;
    N3DS_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void*)sp->k1, 64 * SPLIT_WAVES,
                                                                (size_t)sp->lds1));
    sp->nb1 = (int64_t)g_num_cu * (per_cu < 1 ? 1 : per_cu);
    const int64_t need1 = (n_rays + SPLIT_WAVES - 1) / SPLIT_WAVES;
    if (sp->nb1 > need1) sp->nb1 = need1;
    per_cu = 0;
    N3DS_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void*)pairs_kernel_ptr(cid),
                                                                64 * sp->pw, (size_t)sp->lds2));
    sp->nb2 = (int64_t)g_num_cu * (per_cu < 1 ? 1 : per_cu);
    if (bwd_blocks(sp->kf, sp->bsf, n_rays, &sp->nbf)) return -1;
    const int64_t tot = (int64_t)S * PG;
    // a MODE 2 wavefront takes rays wid, wid + nw, ...: at most rpw of them
    const int64_t nw = sp->nb1 * SPLIT_WAVES, rpw = (n_rays + nw - 1) / nw;
    sp->cap = (rpw * g_tune[TUNE_SPLIT_CAP] + 63) / 64 * 64;
    sp->off_rows = 0;
    sp->off_qf = (sp->nb2 + 1) * tot;
    sp->off_lists = sp->off_qf + sp->nbf * sp->bsf.waves * ring_floats_per_wave();
    sp->off_flags = sp->off_lists + (int64_t)NCLASS * nw * RW * sp->cap;
    sp->total = sp->off_flags + ((int64_t)NCLASS * nw + 1) + 1 + n_rays;  // counts + total, overflow, ray flags
    sp->ok = true;
    return N3DS_OK;
}

// Floats of workspace the split adjoint needs (0: it does not apply); a caller that keeps
// the forward's hits gives the backward max(this, n3ds_render_backward_workspace_floats).
N3DS_EXPORT int64_t n3ds_render_backward_split_floats(int64_t n_rays, int n_strokes, int shape_width,
                                                      int color_width) {
    if (device_info()) return -1;
    if (n_rays <= 0 || n_strokes <= 0) return 0;
    const int PG = shape_width + color_width + 1;
    int cid = -1;
    for (int c = 0; c < 4; c++)
        if (num_color_params(c) == color_width) cid = c;
    if (cid < 0) return 0;
    int64_t need = 0;
    for (int ns = 1; ns <= 4; ns++) {
        SplitPlan sp;
        if (split_plan(ns, cid, n_rays, n_strokes, PG, &sp)) return -1;
        if (!sp.ok) return 0;
        if (sp.total > need) need = sp.total;
    }
    return need;
}

// Number of floats of device workspace the backward needs (per-block partial gradient rows).
N3DS_EXPORT int64_t n3ds_render_backward_workspace_floats(int64_t n_rays, int n_strokes, int shape_width,
                                                          int color_width) {
    if (device_info()) return -1;
    const int64_t PG = shape_width + color_width + 1;
    int cid = 0;
    for (int c = 0; c < 4; c++)
        if (num_color_params(c) == color_width) cid = c;
    // the largest over the sample-count variants: gradient rows (per persistent block when
    // they fit in LDS, else one global set) and the pending-pair queues
    const int64_t tot = (int64_t)n_strokes * PG;
    int64_t need = tot;
    for (int ns = 1; ns <= 4; ns++) {
        const BwdShape bs = bwd_shape(ns, cid, n_strokes, (int)PG);
        if (bs.lds_bytes > g_max_lds) continue;
        int64_t nb = 0;
        if (bwd_blocks(bwd_kernel_ptr(ns, cid, g_tune[TUNE_BWD_QUEUE] != 0 ? 1 : 0), bs, n_rays, &nb)) return -1;
        int64_t n = bs.use_lds ? nb * tot : tot;
        if (g_tune[TUNE_BWD_QUEUE]) n += nb * bs.waves * ring_floats_per_wave();
        if (n > need) need = n;
    }
    return need;
}

// Backward body shared by n3ds_render_backward (caller workspace) and
// n3ds_render_backward_noalloc (the library's fixed scratch).  With fit_blocks
// the persistent block count is lowered until the partial rows fit the
// workspace; the results depend on the block count only through the float
// summation order of the gradients.
static int render_backward_impl(const float* origins, const float* directions, const float* viewdirs,
                                const float* radii, const float* tdist, const float* table, const float* color_params,
                                const float* density_params, int64_t n_rays, int n_samples, int n_strokes,
                                int shape_width, int color_width, int color_id, float sdf_delta,
                                int use_laplace_transform, int inv_scale_radius, int cone, float coord_div,
                                const float* bg_color, const float* bg_rgbs, int opaque_background, float cull_eps, int adjoint_exact, const float* grad_rgb,
                                const float* grad_acc, const float* grad_depth, const float* state, const int* hit_dir,
                                const int* hit_buf, const float* pair_buf, int cap_hits, int cap_pairs,
                                float* workspace, int64_t workspace_floats, bool fit_blocks, float* grad_shape,
                                float* grad_color, float* grad_density, void* stream) {
    int st = render_check(n_rays, n_samples, n_strokes, shape_width, color_width, color_id);
    if (st) return st;
    st = device_info();
    if (st) return st;
    if (n_rays == 0 || n_strokes == 0) return N3DS_OK;
    RenderArgs A = make_args(origins, directions, viewdirs, radii, tdist, table, color_params, density_params,
                             n_rays, n_samples, n_strokes, shape_width, color_width, color_id, sdf_delta,
                             use_laplace_transform, inv_scale_radius, cone, coord_div, bg_color, bg_rgbs, opaque_background, cull_eps,
                             adjoint_exact);
    st = set_keep(A, const_cast<int*>(hit_dir), const_cast<int*>(hit_buf), const_cast<float*>(pair_buf), cap_hits,
                  cap_pairs);
    if (st) return st;
    const int PG = shape_width + color_width + 1;
    const int ns = (n_samples + 63) / 64;
    const BwdShape bs = bwd_shape(ns, color_id, n_strokes, PG);
    N3DS_REQUIRE(bs.lds_bytes <= g_max_lds, "backward needs %lld bytes of LDS per block", (long long)bs.lds_bytes);
    const int64_t tot = (int64_t)n_strokes * PG;
    if (fit_blocks)
        N3DS_REQUIRE(tot <= workspace_floats,
                     "%d strokes x %d gradient columns exceed the %lld-float scratch of the allocation-free backward",
                     n_strokes, PG, (long long)workspace_floats);
    hipStream_t s = (hipStream_t)stream;
    if (!fit_blocks && A.hdir && A.nch <= 64) {  // the split adjoint, when the workspace has room
        SplitPlan sp;
        st = split_plan(ns, color_id, n_rays, n_strokes, PG, &sp);
        if (st) return st;
        if (sp.ok && workspace_floats >= sp.total) {
            float* rows = workspace + sp.off_rows;
            const int64_t nw = sp.nb1 * SPLIT_WAVES;
            A.sp_list = workspace + sp.off_lists;
            A.sp_count = (int*)(workspace + sp.off_flags);
            A.sp_ovf = A.sp_count + NCLASS * nw + 1;
            A.fbflag = A.sp_ovf + 1;
            A.sp_cap = sp.cap;
            A.sp_nw = (int)nw;
            // the overflow flag and the ray flags start clear (every MODE 2 wavefront writes
            // its own counts)
            const int64_t nflag = 1 + n_rays;
            hipLaunchKernelGGL(fill_zero_kernel, dim3((unsigned)((nflag + 255) / 256)), dim3(256), 0, s,
                               (float*)A.sp_ovf, nflag);
            N3DS_LAUNCH_CHECK();
            hipLaunchKernelGGL(sp.k1, dim3((unsigned)sp.nb1), dim3(64 * SPLIT_WAVES), (size_t)sp.lds1, s, A, grad_rgb,
                               grad_acc, grad_depth, state, (float*)nullptr, 0, (float*)nullptr, (float*)nullptr);
            N3DS_LAUNCH_CHECK();
            hipLaunchKernelGGL(group_scan_kernel, dim3(1), dim3(1024), 0, s, A.sp_count, (int64_t)NCLASS * nw,
                               rows + sp.nb2 * tot, tot);
            N3DS_LAUNCH_CHECK();
            hipLaunchKernelGGL(pairs_kernel_ptr(color_id), dim3((unsigned)sp.nb2), dim3(64 * sp.pw), (size_t)sp.lds2,
                               s, A, rows, 1, (float*)nullptr);
            N3DS_LAUNCH_CHECK();
            hipLaunchKernelGGL(sp.kf, dim3((unsigned)sp.nbf), dim3(64 * sp.bsf.waves), (size_t)sp.bsf.lds_bytes, s, A,
                               grad_rgb, grad_acc, grad_depth, state, (float*)nullptr, 0, rows + sp.nb2 * tot,
                               workspace + sp.off_qf);
            N3DS_LAUNCH_CHECK();
            const int nparts = (int)(sp.nb2 + 1);
            hipLaunchKernelGGL(render_reduce_kernel,
                               dim3((unsigned)((tot + 255) / 256), (unsigned)((nparts + RED_PARTS - 1) / RED_PARTS)),
                               dim3(256), 0, s, rows, nparts, n_strokes, shape_width, color_width, grad_shape,
                               grad_color, grad_density);
            N3DS_LAUNCH_CHECK();
            return N3DS_OK;
        }
    }
    // the pending-pair queues follow the gradient rows when the workspace has room for them
    bool queue = !fit_blocks && g_tune[TUNE_BWD_QUEUE] != 0;
    BwdKernel kern = nullptr;
    int64_t nb = 0, ws_need = 0;
    for (;;) {
        kern = bwd_kernel_ptr(ns, color_id, queue ? 1 : 0);
        st = bwd_blocks(kern, bs, n_rays, &nb);
        if (st) return st;
        if (fit_blocks && bs.use_lds && nb * tot > workspace_floats) nb = workspace_floats / tot;
        ws_need = bs.use_lds ? nb * tot : tot;
        if (!queue || workspace_floats >= ws_need + nb * bs.waves * ring_floats_per_wave()) break;
        queue = false;
    }
    N3DS_REQUIRE(workspace_floats >= ws_need, "workspace has %lld floats, need %lld", (long long)workspace_floats,
                 (long long)ws_need);
    float* queues = queue ? workspace + ws_need : nullptr;
    if (!bs.use_lds) {
        hipLaunchKernelGGL(fill_zero_kernel, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, workspace, tot);
        N3DS_LAUNCH_CHECK();
    }
    float* parts = bs.use_lds ? workspace : nullptr;
    float* direct = bs.use_lds ? nullptr : workspace;
    hipLaunchKernelGGL(kern, dim3((unsigned)nb), dim3(64 * bs.waves), (size_t)bs.lds_bytes, s, A, grad_rgb, grad_acc,
                       grad_depth, state, parts, bs.use_lds, direct, queues);
    N3DS_LAUNCH_CHECK();
    const int nparts = bs.use_lds ? (int)nb : 1;
    hipLaunchKernelGGL(render_reduce_kernel,
                       dim3((unsigned)((tot + 255) / 256), (unsigned)((nparts + RED_PARTS - 1) / RED_PARTS)), dim3(256),
                       0, s, workspace,
                       bs.use_lds ? (int)nb : 1, n_strokes, shape_width, color_width, grad_shape, grad_color,
                       grad_density);
    N3DS_LAUNCH_CHECK();
    return N3DS_OK;
}

N3DS_EXPORT int n3ds_render_backward(const float* origins, const float* directions, const float* viewdirs,
                                     const float* radii, const float* tdist, const float* table,
                                     const float* color_params, const float* density_params, int64_t n_rays,
                                     int n_samples, int n_strokes, int shape_width, int color_width, int color_id,
                                     float sdf_delta, int use_laplace_transform, int inv_scale_radius, int cone,
                                     float coord_div, const float* bg_color, const float* bg_rgbs, int opaque_background, float cull_eps, int adjoint_exact,
                                     const float* grad_rgb, const float* grad_acc, const float* grad_depth,
                                     const float* state, const int* hit_dir, const int* hit_buf,
                                     const float* pair_buf, int cap_hits, int cap_pairs, float* workspace,
                                     int64_t workspace_floats, float* grad_shape, float* grad_color,
                                     float* grad_density, void* stream) {
    return render_backward_impl(origins, directions, viewdirs, radii, tdist, table, color_params, density_params,
                                n_rays, n_samples, n_strokes, shape_width, color_width, color_id, sdf_delta,
                                use_laplace_transform, inv_scale_radius, cone, coord_div, bg_color, bg_rgbs, opaque_background, cull_eps,
                                adjoint_exact, grad_rgb, grad_acc, grad_depth, state, hit_dir, hit_buf, pair_buf,
                                cap_hits, cap_pairs, workspace, workspace_floats, false, grad_shape, grad_color,
                                grad_density, stream);
}

// ---------------------------------------------------------------------------
// allocation-free backward: the partial gradient rows live in a fixed
// file-scope device array, one copy per GPU, held until the process ends
// ---------------------------------------------------------------------------
constexpr int64_t NOALLOC_SCRATCH_FLOATS = int64_t(8) << 20;  // 32 MiB
constexpr int NOALLOC_MAX_DEVICES = 64;
static __device__ __align__(256) float scratch[NOALLOC_SCRATCH_FLOATS];
// device address of `scratch` per device, from hipGetSymbolAddress with that device current
static float* g_scratch_dev[NOALLOC_MAX_DEVICES];
static std::mutex g_scratch_mu;

static int scratch_address(float** out) {
    int dev = 0;
    N3DS_HIP_CHECK(hipGetDevice(&dev));
    N3DS_REQUIRE(dev >= 0 && dev < NOALLOC_MAX_DEVICES, "device %d beyond the allocation-free scratch table", dev);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_dev[dev] == nullptr) {
        void* p = nullptr;
        N3DS_HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        N3DS_REQUIRE(p != nullptr, "hipGetSymbolAddress returned no address for the scratch array");
        g_scratch_dev[dev] = static_cast<float*>(p);
    }
    *out = g_scratch_dev[dev];
    return N3DS_OK;
}

N3DS_EXPORT int64_t n3ds_render_backward_noalloc_scratch_floats(void) { return NOALLOC_SCRATCH_FLOATS; }

N3DS_EXPORT int n3ds_render_backward_noalloc(const float* origins, const float* directions, const float* viewdirs,
                                             const float* radii, const float* tdist, const float* table,
                                             const float* color_params, const float* density_params, int64_t n_rays,
                                             int n_samples, int n_strokes, int shape_width, int color_width,
                                             int color_id, float sdf_delta, int use_laplace_transform,
                                             int inv_scale_radius, int cone, float coord_div, const float* bg_color, const float* bg_rgbs, int opaque_background,
                                             float cull_eps, int adjoint_exact, const float* grad_rgb,
                                             const float* grad_acc, const float* grad_depth, const float* state,
                                             const int* hit_dir, const int* hit_buf, const float* pair_buf,
                                             int cap_hits, int cap_pairs, float* grad_shape, float* grad_color,
                                             float* grad_density, void* stream) {
    float* ws = nullptr;
    const int st = scratch_address(&ws);
    if (st) return st;
    return render_backward_impl(origins, directions, viewdirs, radii, tdist, table, color_params, density_params,
                                n_rays, n_samples, n_strokes, shape_width, color_width, color_id, sdf_delta,
                                use_laplace_transform, inv_scale_radius, cone, coord_div, bg_color, bg_rgbs, opaque_background, cull_eps,
                                adjoint_exact, grad_rgb, grad_acc, grad_depth, state, hit_dir, hit_buf, pair_buf,
                                cap_hits, cap_pairs, ws, NOALLOC_SCRATCH_FLOATS, true, grad_shape, grad_color,
                                grad_density, stream);
}

// ---------------------------------------------------------------------------
// batches of independent scenes (one launch per stage for all of them)
// ---------------------------------------------------------------------------

// validate a host scene table (fields 0-3) against the packed sizes; returns the largest scene's strokes
static int batch_check(const int32_t* sc, int n_scenes, int64_t n_rays, int n_strokes, int* s_max) {
    N3DS_REQUIRE(sc && n_scenes >= 1, "a batch needs a scene table with at least one scene");
    N3DS_REQUIRE(n_rays < (int64_t)1 << 31 && n_strokes >= 0, "batched rays / strokes must fit int32");
    int64_t r = 0, k = 0;
    int m = 0;
    for (int i = 0; i < n_scenes; i++) {
        const int32_t* e = sc + (int64_t)i * SCENE_REC;
        N3DS_REQUIRE(e[0] == r && e[1] >= 0 && e[2] == k && e[3] >= 0,
                     "scene %d: rays / strokes must be packed in scene order (got first ray %d, first stroke %d)", i,
                     e[0], e[2]);
        r += e[1];
        k += e[3];
        if (e[3] > m) m = e[3];
    }
    N3DS_REQUIRE(r == n_rays && k == n_strokes, "scene table covers %lld rays / %lld strokes, arrays hold %lld / %d",
                 (long long)r, (long long)k, (long long)n_rays, n_strokes);
    *s_max = m;
    return N3DS_OK;
}

// persistent backward blocks that fit on the device at once for this launch shape
static int bwd_resident_blocks(int ns, int cid, const BwdShape& bs, int64_t* out) {
    BwdKernel kern = bwd_kernel_ptr(ns, cid);
    if (bs.lds_bytes > 65536)
        N3DS_HIP_CHECK(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize,
                                           (int)bs.lds_bytes));
    int per_cu = 0;
    N3DS_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, 64 * bs.waves, (size_t)bs.lds_bytes));
    if (per_cu < 1) per_cu = 1;
    int64_t nb = (int64_t)g_num_cu * per_cu;
    if (nb > (int64_t)g_num_cu * 8) nb = (int64_t)g_num_cu * 8;
    *out = nb;
    return N3DS_OK;
}

// Backward blocks per scene: g_tune[TUNE_BATCH_OVERSUB] times what is resident at
// once, so the hardware dispatcher evens out scenes and rays of uneven cost; each
// block gets at least BATCH_MIN_RAYS rays per wavefront.
constexpr int BATCH_MIN_RAYS = 4;

N3DS_EXPORT int n3ds_render_batch_plan(int32_t* scenes, int n_scenes, int n_samples, int shape_width,
                                       int color_width, int color_id, int32_t* block_map, int block_map_cap) {
    int64_t R = 0;
    int K = 0;
    for (int i = 0; i < n_scenes && scenes; i++) {
        R += scenes[(int64_t)i * SCENE_REC + 1];
        K += scenes[(int64_t)i * SCENE_REC + 3];
    }
    int s_max = 0;
    int st = batch_check(scenes, n_scenes, R, K, &s_max);
    if (st) return -st;
    st = render_check(R, n_samples, s_max, shape_width, color_width, color_id);
    if (st) return -st;
    st = device_info();
    if (st) return -st;
    const int ns = (n_samples + 63) / 64, PG = shape_width + color_width + 1;
    // launch class of each scene: wavefronts per block that fit its LDS accumulators
    // (0: accumulators in global memory); one backward launch per class
    const int classes[4] = {8, 4, 2, 0};
    double work[4] = {0, 0, 0, 0}, tot = 0.0;
    int64_t resident[4] = {0, 0, 0, 0};
    for (int i = 0; i < n_scenes; i++) {
        int32_t* e = scenes + (int64_t)i * SCENE_REC;
        const BwdShape bs = bwd_shape(ns, color_id, e[3], PG);
        e[7] = bs.use_lds ? bs.waves : 0;
        if (e[1] > 0 && e[3] > 0) {
            const double w = (double)e[1] * (e[3] + 32.0);
            tot += w;
            for (int c = 0; c < 4; c++)
                if (classes[c] == e[7]) work[c] += w;
        }
    }
    for (int c = 0; c < 4; c++) {
        if (work[c] == 0.0) continue;
        int sm = 0;
        for (int i = 0; i < n_scenes; i++)
            if (scenes[(int64_t)i * SCENE_REC + 7] == classes[c] && scenes[(int64_t)i * SCENE_REC + 3] > sm)
                sm = scenes[(int64_t)i * SCENE_REC + 3];
        st = bwd_resident_blocks(ns, color_id, bwd_shape(ns, color_id, sm, PG), &resident[c]);
        if (st) return -st;
    }
    // culling tasks: each scene's rays x its own chunks, scene after scene
    {
        int64_t task = 0;
        for (int i = 0; i < n_scenes; i++) {
            int32_t* e = scenes + (int64_t)i * SCENE_REC;
            N3DS_REQUIRE(task < (1ll << 31), "batch too large for the culling plan");
            e[8] = (int32_t)task;
            e[9] = 0;
            const int G = g_tune[TUNE_CULL_GROUP];
            task += (int64_t)e[1] * (((e[3] + 63) / 64 + G - 1) / G);
        }
    }
    // blocks numbered class by class, scene by scene within a class; partial rows packed likewise
    int64_t b = 0, poff = 0;
    for (int c = 0; c < 4; c++) {
        for (int i = 0; i < n_scenes; i++) {
            int32_t* e = scenes + (int64_t)i * SCENE_REC;
            if (e[7] != classes[c]) continue;
            int64_t g = 0;
            if (e[1] > 0 && e[3] > 0) {
                const double w = (double)e[1] * (e[3] + 32.0);
                const int waves = e[7] ? e[7] : BWAVES;
                g = (int64_t)((double)g_tune[TUNE_BATCH_OVERSUB] * resident[c] * w / tot + 0.5);
                const int64_t cap = (e[1] + (int64_t)waves * BATCH_MIN_RAYS - 1) / ((int64_t)waves * BATCH_MIN_RAYS);
                if (g > cap) g = cap;
                if (g < 1) g = 1;
            }
            e[4] = (int32_t)b;
            e[5] = (int32_t)g;
            e[6] = (int32_t)poff;
            if (e[7] > 0) poff += g * e[3] * PG;
            if (block_map && b + g <= block_map_cap)
                for (int64_t k = 0; k < g; k++) block_map[b + k] = i;
            b += g;
            N3DS_REQUIRE(b < (1 << 30) && poff < ((int64_t)1 << 31) - 1, "batch too large for the backward plan");
        }
    }
    return (int)b;
}

N3DS_EXPORT int n3ds_render_batch_forward(const float* origins, const float* directions, const float* viewdirs,
                                          const float* radii, const float* tdist, const float* table,
                                          const float* color_params, const float* density_params,
                                          const int32_t* scenes_host, const int32_t* scenes, int n_scenes,
                                          int64_t n_rays, int n_samples, int n_strokes, int shape_width,
                                          int color_width, int color_id, float sdf_delta, int use_laplace_transform,
                                          int inv_scale_radius, int cone, float coord_div, const float* bg_color, const float* bg_rgbs, int opaque_background,
                                          float cull_eps, float* rgb_out, float* acc_out, float* depth_out,
                                          float* state_out, unsigned long long* counters, int* hit_dir,
                                          int* hit_buf, float* pair_buf, int cap_hits, int cap_pairs,
                                          void* workspace, int64_t workspace_bytes, void* stream) {
    int s_max = 0;
    int st = batch_check(scenes_host, n_scenes, n_rays, n_strokes, &s_max);
    if (st) return st;
    N3DS_REQUIRE(scenes, "the device scene table is missing");
    st = render_check(n_rays, n_samples, s_max, shape_width, color_width, color_id);
    if (st) return st;
    if (n_rays == 0) return N3DS_OK;
    RenderArgs A = make_args(origins, directions, viewdirs, radii, tdist, table, color_params, density_params,
                             n_rays, n_samples, s_max, shape_width, color_width, color_id, sdf_delta,
                             use_laplace_transform, inv_scale_radius, cone, coord_div, bg_color, bg_rgbs, opaque_background, cull_eps, 0);
    A.scenes = scenes;
    A.n_scenes = n_scenes;
    st = set_keep(A, hit_dir, hit_buf, pair_buf, cap_hits, cap_pairs);
    if (st) return st;
    dim3 grid((unsigned)((n_rays + FWAVES - 1) / FWAVES)), block(64 * FWAVES);
    const int ns = (n_samples + 63) / 64;
    int64_t ntask = 0;
    for (int q = 0; q < n_scenes; q++) {
        const int32_t* e = scenes_host + (int64_t)q * SCENE_REC;
        N3DS_REQUIRE(e[8] == ntask, "scene %d has no valid culling plan (run n3ds_render_batch_plan)", q);
        ntask += (int64_t)e[1] * (((e[3] + 63) / 64 + A.cull_group - 1) / A.cull_group);
    }
    N3DS_REQUIRE(ntask < (1ll << 31), "too many (ray, chunk) culling tasks");
    st = launch_cull(A, ns, workspace, workspace_bytes, (hipStream_t)stream, ntask);
    if (st) return st;
    hipLaunchKernelGGL(A.ccounts ? fwd_pre_kernel_ptr(ns, color_id) : fwd_kernel_ptr(ns, color_id), grid, block,
                       fwd_lds_bytes(ns, color_id), (hipStream_t)stream,
                       A, rgb_out, acc_out, depth_out, state_out, counters);
    N3DS_LAUNCH_CHECK();
    return N3DS_OK;
}

// workspace of a planned batch: packed partial rows, then global accumulators for
// all strokes if any scene's accumulators do not fit in LDS
static void batch_workspace(const int32_t* sc, int n_scenes, int PG, int64_t* parts, int64_t* direct) {
    int64_t p = 0, K = 0;
    bool any_direct = false;
    for (int i = 0; i < n_scenes; i++) {
        const int32_t* e = sc + (int64_t)i * SCENE_REC;
        K += e[3];
        if (e[7] > 0) {
            const int64_t end = e[6] + (int64_t)e[5] * e[3] * PG;
            if (end > p) p = end;
        } else if (e[5] > 0) {
            any_direct = true;
        }
    }
    *parts = p;
    *direct = any_direct ? K * PG : 0;
}

N3DS_EXPORT int64_t n3ds_render_batch_backward_workspace_floats(const int32_t* scenes_host, int n_scenes,
                                                                int n_samples, int shape_width, int color_width,
                                                                int color_id) {
    (void)n_samples;
    (void)color_id;
    if (!scenes_host || n_scenes < 1) return -1;
    int64_t p, d;
    batch_workspace(scenes_host, n_scenes, shape_width + color_width + 1, &p, &d);
    return p + d;
}

N3DS_EXPORT int n3ds_render_batch_backward(const float* origins, const float* directions, const float* viewdirs,
                                           const float* radii, const float* tdist, const float* table,
                                           const float* color_params, const float* density_params,
                                           const int32_t* scenes_host, const int32_t* scenes, int n_scenes,
                                           int64_t n_rays, int n_samples, int n_strokes, int shape_width,
                                           int color_width, int color_id, float sdf_delta, int use_laplace_transform,
                                           int inv_scale_radius, int cone, float coord_div, const float* bg_color, const float* bg_rgbs, int opaque_background,
                                           float cull_eps, int adjoint_exact, const float* grad_rgb,
                                           const float* grad_acc, const float* grad_depth, const float* state,
                                           const int* hit_dir, const int* hit_buf, const float* pair_buf,
                                           int cap_hits, int cap_pairs, float* workspace, int64_t workspace_floats,
                                           float* grad_shape, float* grad_color, float* grad_density, void* stream) {
    int s_max = 0;
    int st = batch_check(scenes_host, n_scenes, n_rays, n_strokes, &s_max);
    if (st) return st;
    N3DS_REQUIRE(scenes, "the device scene table is missing");
    st = render_check(n_rays, n_samples, s_max, shape_width, color_width, color_id);
    if (st) return st;
    st = device_info();
    if (st) return st;
    const int PG = shape_width + color_width + 1;
    const int ns = (n_samples + 63) / 64;
    int64_t nb = 0;
    for (int i = 0; i < n_scenes; i++) {
        const int32_t* e = scenes_host + (int64_t)i * SCENE_REC;
        const BwdShape bs = bwd_shape(ns, color_id, e[3], PG);
        N3DS_REQUIRE(e[5] >= 0 && (e[5] > 0 || e[1] == 0 || e[3] == 0) && e[5] <= e[1] &&
                         e[7] == (bs.use_lds ? bs.waves : 0),
                     "scene %d has no valid backward plan (run n3ds_render_batch_plan)", i);
        nb += e[5];
    }
    // the block map (after the scene records, same in the host and device copies)
    const int32_t* map = scenes_host + (int64_t)n_scenes * SCENE_REC;
    for (int i = 0; i < n_scenes; i++) {
        const int32_t* e = scenes_host + (int64_t)i * SCENE_REC;
        N3DS_REQUIRE(e[4] >= 0 && e[4] + (int64_t)e[5] <= nb, "scene %d: backward blocks out of range", i);
        for (int k = 0; k < e[5]; k++)
            N3DS_REQUIRE(map[e[4] + k] == i, "block map entry %d is not scene %d", e[4] + k, i);
    }
    if (n_rays == 0 || n_strokes == 0 || nb == 0) return N3DS_OK;
    int64_t wp, wd;
    batch_workspace(scenes_host, n_scenes, PG, &wp, &wd);
    N3DS_REQUIRE(workspace_floats >= wp + wd, "workspace has %lld floats, need %lld", (long long)workspace_floats,
                 (long long)(wp + wd));
    float* parts = workspace;
    float* direct = workspace + wp;
    hipStream_t s = (hipStream_t)stream;
    if (wd > 0) {
        hipLaunchKernelGGL(fill_zero_kernel, dim3((unsigned)((wd + 255) / 256)), dim3(256), 0, s, direct, wd);
        N3DS_LAUNCH_CHECK();
    }
    // one launch per class of scenes (same wavefronts per block); a class's blocks are consecutive
    const int classes[4] = {8, 4, 2, 0};
    for (int c = 0; c < 4; c++) {
        int64_t b0 = -1, b1 = -1;
        int sm = 0;
        for (int i = 0; i < n_scenes; i++) {
            const int32_t* e = scenes_host + (int64_t)i * SCENE_REC;
            if (e[7] != classes[c] || e[5] == 0) continue;
            if (b0 < 0 || e[4] < b0) b0 = e[4];
            if (e[4] + e[5] > b1) b1 = e[4] + e[5];
            if (e[3] > sm) sm = e[3];
        }
        if (b0 < 0) continue;
        const BwdShape bs = bwd_shape(ns, color_id, sm, PG);
        N3DS_REQUIRE(bs.lds_bytes <= g_max_lds, "backward needs %lld bytes of LDS per block",
                     (long long)bs.lds_bytes);
        RenderArgs A = make_args(origins, directions, viewdirs, radii, tdist, table, color_params, density_params,
                                 n_rays, n_samples, sm, shape_width, color_width, color_id, sdf_delta,
                                 use_laplace_transform, inv_scale_radius, cone, coord_div, bg_color, bg_rgbs, opaque_background, cull_eps,
                                 adjoint_exact);
        A.nch_stride = (s_max + 63) / 64;
        A.scenes = scenes;
        A.n_scenes = n_scenes;
        A.block_base = (int)b0;
        st = set_keep(A, const_cast<int*>(hit_dir), const_cast<int*>(hit_buf), const_cast<float*>(pair_buf),
                      cap_hits, cap_pairs);
        if (st) return st;
        BwdKernel kern = bwd_kernel_ptr(ns, color_id);
        if (bs.lds_bytes > 65536)
            N3DS_HIP_CHECK(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize,
                                               (int)bs.lds_bytes));
        hipLaunchKernelGGL(kern, dim3((unsigned)(b1 - b0)), dim3(64 * bs.waves), (size_t)bs.lds_bytes, s, A,
                           grad_rgb, grad_acc, grad_depth, state, bs.use_lds ? parts : nullptr, bs.use_lds,
                           bs.use_lds ? nullptr : direct, nullptr);
        N3DS_LAUNCH_CHECK();
    }
    const int64_t tot = (int64_t)n_strokes * PG;
    int max_blocks = 1;
    for (int i = 0; i < n_scenes; i++) max_blocks = max(max_blocks, (int)scenes_host[(int64_t)i * SCENE_REC + 5]);
    hipLaunchKernelGGL(render_batch_reduce_kernel,
                       dim3((unsigned)((tot + 255) / 256), (unsigned)((max_blocks + RED_PARTS - 1) / RED_PARTS)),
                       dim3(256), 0, s, parts,
                       direct, scenes, n_scenes, n_strokes, shape_width, color_width, grad_shape, grad_color,
                       grad_density);
    N3DS_LAUNCH_CHECK();
    return N3DS_OK;
}
