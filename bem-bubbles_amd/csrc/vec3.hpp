// Small FP64 3-vector helpers for the per-triangle / per-vertex kernels.
#pragma once
#include <hip/hip_runtime.h>

namespace bemhip {

struct d3 { double x, y, z; };
__device__ inline d3 ld3(const double* p) { return {p[0], p[1], p[2]}; }
__device__ inline d3 sub(d3 a, d3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ inline d3 add(d3 a, d3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ inline d3 scl(double s, d3 a) { return {s * a.x, s * a.y, s * a.z}; }
__device__ inline d3 fma3(double s, d3 a, d3 b) { return {s * a.x + b.x, s * a.y + b.y, s * a.z + b.z}; }
__device__ inline double dot(d3 a, d3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ inline d3 crs(d3 a, d3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
__device__ inline void st3(double* p, d3 v) { p[0] = v.x; p[1] = v.y; p[2] = v.z; }

// 1/sqrt(a) for a > 0: hardware estimate (relative error ~5e-8 on gfx950)
// refined by one Newton step, relative error < 5e-15 (measured over
// [1e-8, 1e8]); callers never pass a = 0 (a quadrature point never coincides
// with a collocation vertex).
__device__ inline double rsqrt_nr(double a) {
    const double y = __builtin_amdgcn_rsq(a);
    const double e = fma(-(a * y), y, 1.0);
    return fma(0.5 * y, e, y);
}

}  // namespace bemhip
