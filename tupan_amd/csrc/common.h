// Shared device helpers for the tupan_amd gfx950 kernels.
//
// Everything here is written for CDNA4 (wave64).  Floating-point code relies
// on IEEE semantics (no fast-math): the reference masks pairs with r2 == 0 by
// a select applied after a NaN may have been produced (tstep_kernel_common.h:
// 50-56), so NaN and Inf must propagate normally.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace tupan {

constexpr int kWave = 64;

// Precision-generic fused multiply-add and max (the __builtin_ forms are
// double-only and would silently promote float arguments).
__device__ __forceinline__ double fma_t(double a, double b, double c) { return __builtin_fma(a, b, c); }
__device__ __forceinline__ float fma_t(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fmax_t(double a, double b) { return __builtin_fmax(a, b); }
__device__ __forceinline__ float fmax_t(float a, float b) { return __builtin_fmaxf(a, b); }

// ---------------------------------------------------------------------------
// Reciprocal square root.
//
// FP64: the hardware estimate v_rsq_f64 refined by one second-order
// Newton-Raphson step on y = x^-1/2.  With e = 1 - x*y0^2 the exact answer is
// y0 * (1 - e)^-1/2 = y0 * (1 + e/2 + 3e^2/8 + ...), so the correction
// y1 = y0 + y0*e*(1/2 + 3e/8) leaves a cubic error term; the residual e is
// formed with a fused multiply-add so it is exact to one rounding of x*y0.
// x == 0 gives Inf/NaN, which callers replace by 0 through the pair mask.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double rsqrt_refined(double x)
{
    double y0 = __builtin_amdgcn_rsq(x);
    double h = x * y0;
    double e = __builtin_fma(-h, y0, 1.0);
    double t = __builtin_fma(e, 0.375, 0.5);
    return __builtin_fma(y0 * e, t, y0);
}

// FP32: v_rsq_f32 refined by one first-order Newton step (the estimate is
// already within about one ulp; the step removes its bias).
__device__ __forceinline__ float rsqrt_refined(float x)
{
    float y0 = __builtin_amdgcn_rsqf(x);
    float h = x * y0;
    float e = __builtin_fmaf(-h, y0, 1.0f);
    return __builtin_fmaf(0.5f * y0, e, y0);
}

// Split an FP64 value into hi + lo floats: hi = fl32(x), lo = fl32(x - hi).
// x - hi is exact in FP64, so hi + lo carries 48 bits of x.
__device__ __forceinline__ void split2(double x, float& hi, float& lo)
{
    hi = static_cast<float>(x);
    lo = static_cast<float>(x - static_cast<double>(hi));
}

// XCD-aware decomposition of a linear workgroup index.  Workgroups are
// dispatched round-robin over the 8 XCDs of an MI355X; taking the j-split
// index as the fast-varying part keeps every workgroup of one XCD on the same
// j range, so each XCD's L2 holds only its share of the j particles.
struct GridPos {
    int iblock;
    int jsplit;
};

__device__ __forceinline__ GridPos grid_pos_of(int bid, int nsplit)
{
    GridPos g;
    g.jsplit = bid % nsplit;
    g.iblock = bid / nsplit;
    return g;
}

__device__ __forceinline__ GridPos grid_pos(int nsplit)
{
    return grid_pos_of(static_cast<int>(blockIdx.x), nsplit);
}

}  // namespace tupan
