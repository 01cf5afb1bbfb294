// Small vector helpers and wavefront primitives for the gfx950 stroke kernels.
//
// A plain struct V3 is used instead of HIP's float3 so operator semantics and
// evaluation order are explicit (the parity tests compare single-precision
// results against a CPU restatement that follows the same order).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace n3ds {

struct V3 {
    float x, y, z;
};

__device__ __forceinline__ V3 v3(float x, float y, float z) { return V3{x, y, z}; }
__device__ __forceinline__ V3 v3s(float s) { return V3{s, s, s}; }
__device__ __forceinline__ V3 operator+(V3 a, V3 b) { return V3{a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ V3 operator-(V3 a, V3 b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ V3 operator-(V3 a) { return V3{-a.x, -a.y, -a.z}; }
__device__ __forceinline__ V3 operator*(V3 a, float s) { return V3{a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ V3 operator*(float s, V3 a) { return V3{s * a.x, s * a.y, s * a.z}; }
__device__ __forceinline__ V3 mul(V3 a, V3 b) { return V3{a.x * b.x, a.y * b.y, a.z * b.z}; }
__device__ __forceinline__ V3 divv(V3 a, V3 b) { return V3{a.x / b.x, a.y / b.y, a.z / b.z}; }
__device__ __forceinline__ float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ V3 cross(V3 a, V3 b) {
    return V3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ V3 vabs(V3 a) { return V3{fabsf(a.x), fabsf(a.y), fabsf(a.z)}; }
__device__ __forceinline__ V3 vmax0(V3 a) { return V3{fmaxf(a.x, 0.f), fmaxf(a.y, 0.f), fmaxf(a.z, 0.f)}; }
// reference sign convention: v < 0 ? -1 : 1
__device__ __forceinline__ float sgn(float v) { return v < 0.f ? -1.f : 1.f; }
__device__ __forceinline__ V3 vsgn(V3 a) { return V3{sgn(a.x), sgn(a.y), sgn(a.z)}; }
// fmaxf(lo, fminf(v, hi)), the clamp of the reference's vector-math header
__device__ __forceinline__ float clampf(float v, float lo, float hi) { return fmaxf(lo, fminf(v, hi)); }
__device__ __forceinline__ float lerpf(float a, float b, float t) { return a + t * (b - a); }
__device__ __forceinline__ V3 ld3(const float* p) { return V3{p[0], p[1], p[2]}; }
__device__ __forceinline__ void add3(float* g, V3 v) { g[0] += v.x; g[1] += v.y; g[2] += v.z; }
// sqrt with zero derivative at exactly zero (vector length without epsilon)
__device__ __forceinline__ float safe_len(V3 v) { float s = dot(v, v); return s > 0.f ? sqrtf(s) : 0.f; }

// ---------------------------------------------------------------------------
// wavefront helpers (64 lanes on CDNA)
// ---------------------------------------------------------------------------
// lane j's value (a scalar register: uniform control flow only, every lane active)
__device__ __forceinline__ int lane_i(int v, int j) { return __builtin_amdgcn_readlane(v, j); }
__device__ __forceinline__ float lane_f(float v, int j) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), j));
}
__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

}  // namespace n3ds
