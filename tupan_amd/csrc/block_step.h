// Device-side block time step shared by the device-driven integrator steps
// (the shared block step of tupan/integrator/__init__.py:48-78).
//
// Control block in device memory: t_curr, t_end, eta, tau, status.
// block_tau_kernel (one workgroup) reduces min |tstep| over the n particles
// ('adaptive' steps) and writes tau: the base step min(|t_end| - |t|, |eta|)
// (get_base_tstep) and, for adaptive steps, the power-of-two block step below
// min |tstep| aligned with t (get_min_block_tstep), in double precision as the
// host does; tau = 0 once t has reached t_end, and once the status is nonzero
// (1: non-finite t, t_end or eta; 2: no positive finite Tstep, a NaN included;
// 3: no aligned block step), which it keeps.  Included by the translation
// units that enqueue device steps; each gets its own copy of the kernel.
#pragma once

#include <float.h>

#include "tupan_internal.h"

namespace tupan {
namespace {

constexpr int kCtlT = 0, kCtlTEnd = 1, kCtlEta = 2, kCtlTau = 3, kCtlStatus = 4;
// Halvings of the block step are bounded: from a step >= 2^-1075 down to
// the spacing of t_next takes at most ~2100.
constexpr int kMaxHalvings = 2200;
constexpr int kTauThreads = 1024;

// min(a, b) that returns NaN when either is NaN, as numpy's min over an
// array with a NaN does (fmin would drop it and pick a step from the rest).
__device__ __forceinline__ double min_nan(double a, double b)
{
    return (a != a || b != b) ? __builtin_nan("") : fmin(a, b);
}

// The work of block_tau_kernel for a workgroup of kTauThreads threads with
// `red` its shared array; thread 0 writes tau (and the status) to ctl.
__device__ __forceinline__ void block_tau_run(int n, const double* ts, int adaptive, double* ctl,
                                              double* red)
{
    const int tid = static_cast<int>(threadIdx.x);
    double lo = DBL_MAX;
    if (adaptive)
        for (int i = tid; i < n; i += kTauThreads) lo = min_nan(lo, fabs(ts[i]));
    red[tid] = lo;
    __syncthreads();
    for (int w = kTauThreads / 2; w > 0; w >>= 1) {
        if (tid < w) red[tid] = min_nan(red[tid], red[tid + w]);
        __syncthreads();
    }
    if (tid != 0) return;
    const double t = ctl[kCtlT], tend = ctl[kCtlTEnd], eta = ctl[kCtlEta];
    if (ctl[kCtlStatus] != 0 || !isfinite(t) || !isfinite(tend) || !isfinite(eta)) {
        ctl[kCtlTau] = 0.0;
        if (ctl[kCtlStatus] == 0) ctl[kCtlStatus] = 1.0;
        return;
    }
    if ((eta > 0 && t >= tend) || (eta < 0 && t <= tend) || eta == 0) {
        ctl[kCtlTau] = 0.0;
        return;
    }
    // integrator/__init__.py:48-57
    double tau = fmin(fabs(tend) - fabs(t), fabs(eta));
    tau = fmax(tau, fabs(tend) * (2 * DBL_EPSILON));
    tau = copysign(tau, eta);
    if (adaptive) {
        // integrator/__init__.py:59-78
        const double min_ts = red[0];
        if (!(min_ts > 0) || !isfinite(min_ts)) {
            // Tstep gave no positive finite step (NaN or zero inputs)
            ctl[kCtlTau] = 0.0;
            ctl[kCtlStatus] = 2.0;
            return;
        }
        const double p = log2(min_ts) - 1;
        if (p < -1070.0 || p > 1020.0) {
            ctl[kCtlTau] = 0.0;
            ctl[kCtlStatus] = 2.0;
            return;
        }
        double bts = ldexp(1.0, static_cast<int>(p));
        const double t_next = t + bts;
        int halvings = 0;
        while (fmod(t_next, bts) != 0 && halvings < kMaxHalvings) {
            bts /= 2;
            ++halvings;
        }
        if (halvings == kMaxHalvings || bts == 0) {
            ctl[kCtlTau] = 0.0;
            ctl[kCtlStatus] = 3.0;
            return;
        }
        bts = copysign(bts, tau);
        if (fabs(bts) > fabs(tau)) bts = tau;
        tau = bts;
    }
    ctl[kCtlTau] = tau;
}

__global__ __launch_bounds__(kTauThreads) void block_tau_kernel(int n, const double* ts,
                                                                   int adaptive, double* ctl)
{
    __shared__ double red[kTauThreads];
    block_tau_run(n, ts, adaptive, ctl, red);
}

}  // namespace
}  // namespace tupan
