// Interval resampling from a step function: the proposal levels' sampler,
// drop-in for source/utils/stepfun.py:sample_intervals (:280-317) with
// sample (:216-254), invert_cdf (:195-202), integrate_weights (:149-170) and
// sorted_interp (:5-24), as called at models.py:168-181.
//
// Mapping: one wavefront per ray.  The ray's bin edges t [m+1], CDF
// cw0 [m+1] and centres [n] are staged in the wave's LDS slice: softmax of the logits by two
// wave reductions (max, sum of exp), the CDF by a 64-lane inclusive scan with
// a running carry, clamped to 1 and closed with exact 0 and 1.  Each lane then
// evaluates centres: the inverse CDF at a sample u is an upper-bound binary
// search in cw0 (cw0 is non-decreasing, so "last entry <= u" is what
// sorted_interp's masked max / min select for sorted t) and one linear
// interpolation; fencepost o takes the midpoint of centres o-1 and o, and the
// first / last the reflection of the end midpoint clamped to the domain.
// Per ray the work is O(m + n log m) and the bytes are the inputs and outputs
// once (HBM / latency bound).
#include <algorithm>

#include "n3ds_capi.h"
#include "n3ds_math.h"

namespace n3ds {

constexpr int SF_MAX_FLOATS = 16384;  // LDS floats per ray slice: 2 (m + 1) + n (<= 64 KB per block)

// LDS ordering between the lanes of one wavefront (no block barrier)
__device__ __forceinline__ void sf_wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ float sf_wave_max(float v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = fmaxf(v, __shfl_xor(v, off, 64));
    return v;
}

// the inverse CDF at u (sorted_interp(u, cw0, t) for non-decreasing cw0, sorted t)
__device__ __forceinline__ float sf_center(const float* cw, const float* te, int m, float u) {
    int lo = 0, hi = m + 1;  // first index with cw[idx] > u, in [1, m] after the clamp
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (cw[mid] <= u) lo = mid + 1;
        else hi = mid;
    }
    const int k = min(max(lo, 1), m);
    const float x0 = cw[k - 1], x1 = cw[k], f0 = te[k - 1], f1 = te[k];
    float off = (u - x0) / (x1 - x0);
    if (off != off) off = 0.f;  // nan_to_num(., 0)
    off = fminf(fmaxf(off, 0.f), 1.f);
    return __fadd_rn(f0, __fmul_rn(off, __fsub_rn(f1, f0)));
}

__global__ __launch_bounds__(256) void sample_intervals_kernel(float* __restrict__ t_out,
                                                               const float* __restrict__ t,
                                                               const float* __restrict__ logits,
                                                               const float* __restrict__ u_base,
                                                               const float* __restrict__ jitter, int jitter_width,
                                                               float max_jitter, int64_t n_rays, int m, int n,
                                                               float minval, float maxval, int waves) {
    extern __shared__ float sf_lds[];
    const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // uniform per wavefront
    float* cw = sf_lds + (int64_t)wave * (2 * (m + 1) + n);
    float* te = cw + (m + 1);
    float* cs = te + (m + 1);
    for (int64_t r = (int64_t)blockIdx.x * waves + wave; r < n_rays; r += (int64_t)gridDim.x * waves) {
        const float* tr = t + r * (m + 1);
        const float* lr = logits + r * m;
        // softmax over the m bins
        float mx = -INFINITY;
        for (int j = lane; j < m; j += 64) mx = fmaxf(mx, lr[j]);
        mx = sf_wave_max(mx);
        float se = 0.f;
        for (int j = lane; j < m; j += 64) se += expf(lr[j] - mx);
        se = wave_sum(se);
        // cw0 = [0, min(cumsum(w[:-1]), 1), 1]
        float carry = 0.f;
        for (int j0 = 0; j0 < m - 1; j0 += 64) {
            const int j = j0 + lane;
            float v = j < m - 1 ? expf(lr[j] - mx) / se : 0.f;
#pragma unroll
            for (int off = 1; off < 64; off <<= 1) {
                const float o = __shfl_up(v, off, 64);
                if (lane >= off) v += o;
            }
            v += carry;
            if (j < m - 1) cw[j + 1] = fminf(v, 1.f);
            carry = __shfl(v, 63, 64);
        }
        if (lane == 0) {
            cw[0] = 0.f;
            cw[m] = 1.f;
        }
        for (int j = lane; j <= m; j += 64) te[j] = tr[j];
        sf_wave_sync();
        float* outr = t_out + r * (n + 1);
        const float* jr = jitter ? jitter + r * jitter_width : nullptr;
        for (int i = lane; i < n; i += 64) {
            float u = u_base[i];
            if (jr) u = __fadd_rn(u, __fmul_rn(jr[jitter_width == 1 ? 0 : i], max_jitter));
            cs[i] = sf_center(cw, te, m, u);
        }
        sf_wave_sync();
        for (int o = lane; o <= n; o += 64) {
            float val;
            if (o == 0 || o == n) {
                const int e = o == 0 ? 0 : n - 1, nb = o == 0 ? 1 : n - 2;  // end centre and its neighbour
                const float ce = cs[e], cn = cs[nb];
                const float midv = (o == 0 ? __fadd_rn(cn, ce) : __fadd_rn(ce, cn)) / 2.f;
                const float refl = __fsub_rn(2.f * ce, midv);
                // clamp_min / clamp_max keep NaN
                val = o == 0 ? (refl < minval ? minval : refl) : (refl > maxval ? maxval : refl);
            } else {
                val = __fadd_rn(cs[o], cs[o - 1]) / 2.f;
            }
            outr[o] = val;
        }
        sf_wave_sync();  // the next ray overwrites the slice
    }
}

// number of entries of the sorted a[0..n) below x (strictly, or at most x when le)
__device__ __forceinline__ int sf_rank(const float* a, int n, float x, bool le) {
    int lo = 0, hi = n;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (le ? a[mid] <= x : a[mid] < x) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// max_dilate_weights (stepfun.py:136-146 over max_dilate :120-133, weight_to_pdf :109-112,
// pdf_to_weight :115-117): per ray, the edges t, t[:-1] - d and t[1:] + d merged in
// sorted order (three sorted runs: each element's place is its index plus its rank in
// the other two runs, ties broken run by run), clipped to the domain; the PDF at edge i
// is the max of p_j over the bins with t_j - d <= e_i < t_{j+1} + d (a contiguous range
// of j for sorted t, 0 when some bin does not cover it); back to weights, renormalised.
// LDS per wave: t [m+1], p [m], merged edges [3m+1].
__global__ __launch_bounds__(256) void max_dilate_kernel(float* __restrict__ t_out, float* __restrict__ w_out,
                                                         const float* __restrict__ t, const float* __restrict__ w,
                                                         int64_t n_rays, int m, float dilation, float minval,
                                                         float maxval, int renormalize, int waves) {
    extern __shared__ float sf_lds[];
    const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // uniform per wavefront
    const int ne = 3 * m + 1;
    float* ts = sf_lds + (int64_t)wave * (5 * m + 2);
    float* ps = ts + (m + 1);
    float* es = ps + m;
    const float eps = 1.1920928955078125e-7f;  // float32 machine epsilon
    for (int64_t r = (int64_t)blockIdx.x * waves + wave; r < n_rays; r += (int64_t)gridDim.x * waves) {
        const float* tr = t + r * (m + 1);
        const float* wr = w + r * m;
        for (int j = lane; j <= m; j += 64) ts[j] = tr[j];
        sf_wave_sync();
        for (int j = lane; j < m; j += 64) ps[j] = wr[j] / fmaxf(__fsub_rn(ts[j + 1], ts[j]), eps);
        // merged edge positions: run A = t [m+1], B = t[:-1] - d [m], C = t[1:] + d [m]
        for (int k = lane; k < ne; k += 64) {
            float v;
            int pos;
            if (k <= m) {  // A: after B / C entries strictly below
                v = ts[k];
                int lo = 0, hi = m;  // B below v: t_j - d < v
                while (lo < hi) { const int mid = (lo + hi) >> 1; if (__fsub_rn(ts[mid], dilation) < v) lo = mid + 1; else hi = mid; }
                int cb = lo;
                lo = 0; hi = m;      // C below v: t_{j+1} + d < v
                while (lo < hi) { const int mid = (lo + hi) >> 1; if (__fadd_rn(ts[mid + 1], dilation) < v) lo = mid + 1; else hi = mid; }
                pos = k + cb + lo;
            } else if (k <= 2 * m) {  // B: after A entries <= v and C entries < v
                const int j = k - (m + 1);
                v = __fsub_rn(ts[j], dilation);
                const int ca = sf_rank(ts, m + 1, v, true);
                int lo = 0, hi = m;
                while (lo < hi) { const int mid = (lo + hi) >> 1; if (__fadd_rn(ts[mid + 1], dilation) < v) lo = mid + 1; else hi = mid; }
                pos = j + ca + lo;
            } else {  // C: after A and B entries <= v
                const int j = k - (2 * m + 1);
                v = __fadd_rn(ts[j + 1], dilation);
                const int ca = sf_rank(ts, m + 1, v, true);
                int lo = 0, hi = m;
                while (lo < hi) { const int mid = (lo + hi) >> 1; if (__fsub_rn(ts[mid], dilation) <= v) lo = mid + 1; else hi = mid; }
                pos = j + ca + lo;
            }
            es[pos] = fminf(fmaxf(v, minval), maxval);  // torch.clip(min, max)
        }
        sf_wave_sync();
        float* tout = t_out + r * ne;
        float* wout = w_out + r * (ne - 1);
        float wsum = 0.f;
        for (int i = lane; i < ne; i += 64) {
            const float e = es[i];
            tout[i] = e;
            if (i == ne - 1) continue;
            // bins covering e: t_j - d <= e (a prefix j < a) and t_{j+1} + d > e (a suffix j >= b)
            int lo = 0, hi = m;
            while (lo < hi) { const int mid = (lo + hi) >> 1; if (__fsub_rn(ts[mid], dilation) <= e) lo = mid + 1; else hi = mid; }
            const int a = lo;
            lo = 0; hi = m;
            while (lo < hi) { const int mid = (lo + hi) >> 1; if (__fadd_rn(ts[mid + 1], dilation) <= e) lo = mid + 1; else hi = mid; }
            const int b = lo;
            float pd = (b == 0 && a == m) ? -INFINITY : 0.f;
            for (int j = b; j < a; j++) pd = fmaxf(pd, ps[j]);
            const float wd = __fmul_rn(pd, __fsub_rn(es[i + 1], e));
            wout[i] = wd;
            wsum += wd;
        }
        if (renormalize) {
            wsum = fmaxf(wave_sum(wsum), eps);
            for (int i = lane; i < ne - 1; i += 64) wout[i] = wout[i] / wsum;
        }
        sf_wave_sync();  // the next ray overwrites the slice
    }
}

}  // namespace n3ds

using namespace n3ds;

N3DS_EXPORT int n3ds_sample_intervals(float* t_out, const float* t, const float* w_logits, const float* u_base,
                                      const float* jitter, int jitter_width, float max_jitter, int64_t n_rays,
                                      int n_bins, int n_samples, float minval, float maxval, void* stream) {
    N3DS_REQUIRE(n_samples > 1, "num_samples must be > 1, is %d.", n_samples);
    N3DS_REQUIRE(n_bins >= 1 && 2 * ((int64_t)n_bins + 1) + n_samples <= SF_MAX_FLOATS,
                 "2 (n_bins + 1) + n_samples = %lld exceeds %d", 2 * ((long long)n_bins + 1) + n_samples, SF_MAX_FLOATS);
    N3DS_REQUIRE(n_rays >= 0, "negative size");
    N3DS_REQUIRE(!jitter || jitter_width == 1 || jitter_width == n_samples,
                 "jitter width must be 1 or n_samples (got %d)", jitter_width);
    if (n_rays == 0) return N3DS_OK;
    N3DS_REQUIRE(t_out && t && w_logits && u_base, "null argument");
    const int per_wave = 2 * (n_bins + 1) + n_samples;
    const int waves = per_wave <= SF_MAX_FLOATS / 4 ? 4 : 1;
    const size_t lds = (size_t)waves * per_wave * sizeof(float);
    const int64_t blocks = std::min<int64_t>((n_rays + waves - 1) / waves, 256 * 32);
    hipLaunchKernelGGL(sample_intervals_kernel, dim3((unsigned)blocks), dim3(64 * waves), lds, (hipStream_t)stream,
                       t_out, t, w_logits, u_base, jitter, jitter_width, max_jitter, n_rays, n_bins, n_samples,
                       minval, maxval, waves);
    N3DS_LAUNCH_CHECK();
    return N3DS_OK;
}

N3DS_EXPORT int n3ds_max_dilate_weights(float* t_out, float* w_out, const float* t, const float* w, int64_t n_rays,
                                        int n_bins, float dilation, float minval, float maxval, int renormalize,
                                        void* stream) {
    N3DS_REQUIRE(n_bins >= 1 && 5 * (int64_t)n_bins + 2 <= 16384, "n_bins %d outside [1, 3276]", n_bins);
    N3DS_REQUIRE(n_rays >= 0, "negative size");
    if (n_rays == 0) return N3DS_OK;
    N3DS_REQUIRE(t_out && w_out && t && w, "null argument");
    const int per_wave = 5 * n_bins + 2;
    const int waves = per_wave <= 4096 ? 4 : 1;
    const size_t lds = (size_t)waves * per_wave * sizeof(float);
    const int64_t blocks = std::min<int64_t>((n_rays + waves - 1) / waves, 256 * 32);
    hipLaunchKernelGGL(max_dilate_kernel, dim3((unsigned)blocks), dim3(64 * waves), lds, (hipStream_t)stream, t_out,
                       w_out, t, w, n_rays, n_bins, dilation, minval, maxval, renormalize, waves);
    N3DS_LAUNCH_CHECK();
    return N3DS_OK;
}
