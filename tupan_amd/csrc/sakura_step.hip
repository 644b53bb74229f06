// Device-driven Sakura steps (tupan/integrator/sakura.py:22-50, 100-117):
// one call enqueues a whole step with no host synchronisation, the step size
// read from and written to a control block in device memory, so a sequence
// of calls can be captured as one HIP graph.  Per step:
//
//   1. 'asakura': Tstep (eta) into scratch arrays;
//   2. sakura_tau_kernel: the base step min(|t_end| - |t|, |eta|) and, for
//      'asakura', eta / sqrt(1 + max_i(w2_a - w2_b)) (w2 = (eta / tstep)^2
//      as the host driver forms it) quantised to the power-of-two block step
//      aligned with t; tau = 0 once t has reached t_end or once the control
//      block's status is nonzero (1: non-finite t, t_end or eta; 2: no
//      finite Tstep maximum; 3: no aligned block step), which it keeps;
//   3. drift by tau/2, the pair drift with flag -1 over tau/2 (two-pass form
//      without host synchronisation, dt read from the control block), apply,
//      the same with flag +1, drift by tau/2;
//   4. sakura_end_kernel: tstep = tau, time += tau, nstep += 1, tstepij from
//      the scratch Tstep output ('asakura'), t += tau.
// Every elementwise kernel does nothing when tau = 0, so a step past t_end
// changes no particle array.  The drifts and updates are the host driver's
// torch operations one rounding at a time (this file is compiled with
// -ffp-contract=off), so device steps equal host-driven ones bit for bit.
#include <float.h>

#include "tupan_internal.h"

namespace tupan {

namespace {

constexpr int kT = 0, kTEnd = 1, kEta = 2, kTau = 3, kStatus = 4, kDt = 5;
constexpr int kMaxHalvings = 2200;
constexpr int kTauThreads = 1024;

struct SakuraStepArrays {
    double *r[3], *v[3], *d[6];
    double *tstep, *tstepij, *time;
    long long* nstep;
    const double* tsij;  // scratch Tstep output
};

__global__ __launch_bounds__(kTauThreads) void sakura_tau_kernel(int n, const double* ts,
                                                                 const double* tsij,
                                                                 int adaptive, double* ctl)
{
    __shared__ double red[kTauThreads];
    __shared__ int bad[kTauThreads];
    const int tid = static_cast<int>(threadIdx.x);
    const double eta_h = ctl[kEta];
    double hi = -DBL_MAX;
    int nan = 0;
    if (adaptive) {
        for (int i = tid; i < n; i += kTauThreads) {
            // sakura.py:100-111 as torch evaluates it: eta / x is
            // reciprocal(x) * eta, x**2 is x * x
            const double a = (1.0 / ts[i]) * eta_h;
            const double b = (1.0 / tsij[i]) * eta_h;
            const double d = a * a - b * b;
            if (d != d) nan = 1;
            hi = fmax(hi, d);
        }
    }
    red[tid] = hi;
    bad[tid] = nan;
    __syncthreads();
    for (int w = kTauThreads / 2; w > 0; w >>= 1) {
        if (tid < w) {
            red[tid] = fmax(red[tid], red[tid + w]);
            bad[tid] |= bad[tid + w];
        }
        __syncthreads();
    }
    if (tid != 0) return;
    const double t = ctl[kT], tend = ctl[kTEnd], eta = ctl[kEta];
    if (ctl[kStatus] != 0 || !isfinite(t) || !isfinite(tend) || !isfinite(eta)) {
        ctl[kTau] = 0.0;
        ctl[kDt] = 0.0;
        if (ctl[kStatus] == 0) ctl[kStatus] = 1.0;
        return;
    }
    if ((eta > 0 && t >= tend) || (eta < 0 && t <= tend) || eta == 0) {
        ctl[kTau] = 0.0;
        ctl[kDt] = 0.0;
        return;
    }
    // integrator/__init__.py:48-57
    double tau = fmin(fabs(tend) - fabs(t), fabs(eta));
    tau = fmax(tau, fabs(tend) * (2 * DBL_EPSILON));
    tau = copysign(tau, eta);
    if (adaptive) {
        const double w2 = red[0];
        const double dt_sakura = eta / sqrt(1 + w2);
        const double min_ts = fabs(dt_sakura);
        if (bad[0] || !(min_ts > 0) || !isfinite(min_ts)) {
            ctl[kTau] = 0.0;
            ctl[kDt] = 0.0;
            ctl[kStatus] = 2.0;
            return;
        }
        // integrator/__init__.py:59-78
        const double p = log2(min_ts) - 1;
        if (p < -1070.0 || p > 1020.0) {
            ctl[kTau] = 0.0;
            ctl[kDt] = 0.0;
            ctl[kStatus] = 2.0;
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
            ctl[kTau] = 0.0;
            ctl[kDt] = 0.0;
            ctl[kStatus] = 3.0;
            return;
        }
        bts = copysign(bts, tau);
        if (fabs(bts) > fabs(tau)) bts = tau;
        tau = bts;
    }
    ctl[kTau] = tau;
    ctl[kDt] = tau / 2;
}

// r += v * tau / 2 (sakura.py:23-25, 47-49)
__global__ __launch_bounds__(256) void sakura_drift_kernel(int n, SakuraStepArrays h,
                                                           const double* ctl)
{
    const int i = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    const double tau = ctl[kTau];
    if (i >= n || tau == 0) return;
#pragma unroll
    for (int d = 0; d < 3; ++d) h.r[d][i] += h.v[d][i] * tau / 2;
}

// r += dr, v += dv (sakura.py:28-33, 36-41)
__global__ __launch_bounds__(256) void sakura_apply_kernel(int n, SakuraStepArrays h,
                                                           const double* ctl)
{
    const int i = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= n || ctl[kTau] == 0) return;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        h.r[d][i] += h.d[d][i];
        h.v[d][i] += h.d[3 + d][i];
    }
}

// sakura.py:113-117 bookkeeping, then t += tau.
__global__ __launch_bounds__(256) void sakura_end_kernel(int n, SakuraStepArrays h,
                                                         int adaptive, double* ctl)
{
    const int i = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    const double tau = ctl[kTau];
    if (tau == 0) return;
    if (i < n) {
        h.tstep[i] = tau;
        h.time[i] += tau;
        h.nstep[i] += 1;
        if (adaptive) h.tstepij[i] = h.tsij[i];
    }
    if (i == 0) ctl[kT] += tau;
}

}  // namespace

int sakura_step_call(int adaptive, int mode, size_t n, void* const* arrays, double* scratch,
                     double* ctl, double eta, int capacity, unsigned long long* counters,
                     int* status, void* stream)
{
    if (n == 0) return TUPAN_OK;
    if (n > static_cast<size_t>(INT32_MAX)) return set_error(TUPAN_ERR_ARG, "too many particles");
    if (!arrays || !ctl || !status || (adaptive && !scratch))
        return set_error(TUPAN_ERR_ARG, "null array, control block or status");
    for (int k = 0; k < TUPAN_S_NROLES; ++k)
        if (!arrays[k]) return set_error(TUPAN_ERR_ARG, "null particle array");
    hipStream_t s = static_cast<hipStream_t>(stream);
    const int ni = static_cast<int>(n);
    auto arr = [&](int role) { return static_cast<double*>(arrays[role]); };
    const double* in8[8] = {arr(TUPAN_S_MASS), arr(TUPAN_S_RX), arr(TUPAN_S_RX + 1),
                            arr(TUPAN_S_RX + 2), arr(TUPAN_S_EPS2), arr(TUPAN_S_VX),
                            arr(TUPAN_S_VX + 1), arr(TUPAN_S_VX + 2)};
    double* out6[6];
    SakuraStepArrays h{};
    for (int d = 0; d < 3; ++d) {
        h.r[d] = arr(TUPAN_S_RX + d);
        h.v[d] = arr(TUPAN_S_VX + d);
    }
    for (int k = 0; k < 6; ++k) h.d[k] = out6[k] = arr(TUPAN_S_DRX + k);
    h.tstep = arr(TUPAN_S_TSTEP);
    h.tstepij = arr(TUPAN_S_TSTEPIJ);
    h.time = arr(TUPAN_S_TIME);
    h.nstep = static_cast<long long*>(arrays[TUPAN_S_NSTEP]);
    double* ts = scratch;
    double* tsij = scratch ? scratch + n : nullptr;
    h.tsij = tsij;
    int rc = TUPAN_OK;
    if (adaptive) {
        double* tout[2] = {ts, tsij};
        rc = pairwise_call<double>(K_TSTEP, mode, n, in8, n, in8, tout, eta, stream);
        if (rc != TUPAN_OK) return rc;
    }
    hipLaunchKernelGGL(sakura_tau_kernel, dim3(1), dim3(kTauThreads), 0, s, ni, ts, tsij,
                       adaptive, ctl);
    const unsigned g = static_cast<unsigned>((n + 255) / 256);
    hipLaunchKernelGGL(sakura_drift_kernel, dim3(g), dim3(256), 0, s, ni, h, ctl);
    for (long flag : {-1L, 1L}) {
        SakuraOptions opt;
        opt.path = 2;
        opt.graph_capacity = capacity;
        opt.dt_dev = ctl + kDt;
        rc = sakura_call_ex(n, in8, n, in8, 0.0, flag, out6, opt, counters, status, stream);
        if (rc != TUPAN_OK) return rc;
        hipLaunchKernelGGL(sakura_apply_kernel, dim3(g), dim3(256), 0, s, ni, h, ctl);
    }
    hipLaunchKernelGGL(sakura_drift_kernel, dim3(g), dim3(256), 0, s, ni, h, ctl);
    hipLaunchKernelGGL(sakura_end_kernel, dim3(g), dim3(256), 0, s, ni, h, adaptive, ctl);
    return check_launch("sakura step");
}

}  // namespace tupan
