// Device-resident SIA recursion for small particle sets.
//
// The hierarchical SIA step (tupan/integrator/sia.py:1089-1123, the 'h'
// variants) visits every level of a subset's slow/fast hierarchy from the
// host: per visit a Tstep call, a split, drifts, kicks (Acc calls) and the
// slow<->fast coupling.  Below a few dozen particles each visit is a handful
// of tiny launches plus host work.  sia_subtree_kernel runs the whole
// recursion of such a subset in one launch of one wave: lane k holds particle
// k of the subset's range, the recursion is an explicit stack of frames, and
// every operation is the host driver's, in the host driver's order:
//
//   recurse(S, tau):  Tstep(S, S); stable partition S -> [slow | fast] with
//                     slow = |tstep| > |tau| (n <= 2: all slow); for each
//                     (op, c) of the bridge: op 0 -> evolve(slow, c tau) then
//                     recurse(fast, c tau); op 1 -> slow<->fast kicks;
//                     t_curr += tau if fast is empty; slow.tstep = tau,
//                     slow.time += tau, slow.nstep += 1
//   evolve(P, dt):    n <= 2: drift (n = 1) or exact two-body (n = 2);
//                     else for each (op, c): op 0 drift(c dt), op 1 kick
//                     (Acc(P, P), v += a c dt)
//
// The pair interactions are ops.h's, evaluated as the tile engine evaluates
// them for sets that fit one tile (one j split, j in order, one flush, frame
// at the set's first i-particle), so the subtree produces the host driver's
// numbers bit for bit.  Drifts and kicks are x + (v * dt), unfused, as the
// host's elementwise updates are.
//
// Compiled with -ffp-contract=off.
#include "block_step.h"
#include "kepler.h"
#include "ops.h"
#include "tupan_internal.h"

namespace tupan {

constexpr int kSubMax = 64;       // particles (one wave)
constexpr int kSubDepth = 64;     // recursion levels
constexpr int kSubOps = 32;       // entries of a bridge or evolve sequence
constexpr int kSubArrays = 32;    // 8-byte arrays permuted with the particles
// Level visits one launch may make.  A hierarchy of L levels below the
// subset visits up to 2^L - 1 of them (sia21), so this admits 16 fully
// populated levels; a set whose steps never exceed the level's tau (NaN
// steps, for one) would otherwise recurse without end.  Past it the kernel
// stops with status 3.
constexpr int kSubVisits = 1 << 16;

enum SubRole {
    R_MASS, R_E2, R_RX, R_RY, R_RZ, R_VX, R_VY, R_VZ,
    R_TIME, R_NSTEP, R_TSTEP, R_TSTEPIJ, R_AX, R_AY, R_AZ, R_NROLES
};

struct SubtreeArgs {
    unsigned long long* arr[kSubArrays];
    int narr;
    int role[R_NROLES];
    int n;
    int acc_mode, tstep_mode;
    double eta, tau;
    int nbridge, nevolve;
    int bridge_op[kSubOps];
    double bridge_c[kSubOps];
    int evolve_op[kSubOps];
    double evolve_c[kSubOps];
    double* t_curr;                // device scalar, updated in place
    unsigned long long* pairs;     // device [2]: Tstep and Acc interactions
    int* status;                   // 1: Kepler failure, 2: stack overflow,
                                   // 3: visit budget exhausted
};

struct Frame {
    int lo, n, ns, pc;
    double tau;
};

__device__ __forceinline__ double* field(const SubtreeArgs& a, int role)
{
    return reinterpret_cast<double*>(a.arr[a.role[role]]);
}

// One pairwise evaluation for i in [ilo, ilo + ni) against j in
// [jlo, jlo + nj), one i per lane, written to the Op's outputs.
template <class Op>
__device__ void pair_op(const SubtreeArgs& a, int ilo, int ni, int jlo, int nj, double eta,
                        double* o0, double* o1, double* o2)
{
    typename Op::Args pa{};
    const int roles[8] = {R_MASS, R_RX, R_RY, R_RZ, R_E2, R_VX, R_VY, R_VZ};
#pragma unroll
    for (int f = 0; f < 8; ++f) {
        pa.ip[f] = field(a, roles[f]) + ilo;
        pa.jp[f] = field(a, roles[f]) + jlo;
    }
    pa.out[0] = o0 + ilo;
    pa.out[1] = o1 ? o1 + ilo : nullptr;
    pa.out[2] = o2 ? o2 + ilo : nullptr;
    pa.ni = ni;
    pa.nj = nj;
    pa.nsplit = 1;
    pa.jchunk = nj;
    pa.eta = eta;
    const int lane = static_cast<int>(threadIdx.x);
    const typename Op::Frame fr = Op::frame(pa, 0);
    if (lane < ni) {
        typename Op::IReg r;
        typename Op::Acc acc;
        Op::load_i(pa, fr, lane, r);
        Op::zero(acc);
        for (int j = 0; j < nj; ++j) {
            typename Op::Slot s;
            Op::load_j(pa, fr, j, s);
            Op::interact(r, s, acc);
        }
        Op::flush(acc);
        double tot[Op::kNacc];
#pragma unroll
        for (int k = 0; k < Op::kNacc; ++k) tot[k] = Op::get(acc, k);
        if constexpr (HasFallback<Op>::value) {
            if (!all_finite(tot)) fallback64<typename Op::Fallback64>(pa, lane, 0, nj, 1, tot);
        }
        Op::store(pa, lane, tot);
    }
    __syncthreads();
}

__device__ void acc_op(const SubtreeArgs& a, int ilo, int ni, int jlo, int nj)
{
    if (a.acc_mode == TUPAN_MODE_SPLIT32) {
        pair_op<OpSplitAcc>(a, ilo, ni, jlo, nj, 0.0, field(a, R_AX), field(a, R_AY),
                            field(a, R_AZ));
    } else {
        pair_op<OpAcc<Fast64>>(a, ilo, ni, jlo, nj, 0.0, field(a, R_AX), field(a, R_AY),
                               field(a, R_AZ));
    }
    if (threadIdx.x == 0) atomicAdd(a.pairs + 1, static_cast<unsigned long long>(ni) * nj);
}

__device__ void tstep_op(const SubtreeArgs& a, int lo, int n)
{
    if (a.tstep_mode == TUPAN_MODE_SPLIT32) {
        pair_op<OpSplitTstep>(a, lo, n, lo, n, a.eta, field(a, R_TSTEP), field(a, R_TSTEPIJ),
                              nullptr);
    } else {
        pair_op<OpTstep<Fast64>>(a, lo, n, lo, n, a.eta, field(a, R_TSTEP),
                                 field(a, R_TSTEPIJ), nullptr);
    }
    if (threadIdx.x == 0) atomicAdd(a.pairs + 0, static_cast<unsigned long long>(n) * n);
}

// x += v * dt for the lanes of [lo, lo + n), v from roles vr.. (position
// drift: vr = R_VX; kick: vr = R_AX onto R_VX).
__device__ void axpy3(const SubtreeArgs& a, int lo, int n, int xr, int vr, double dt)
{
    const int lane = static_cast<int>(threadIdx.x);
    if (lane < n) {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            double* x = field(a, xr + c) + lo + lane;
            const double v = field(a, vr + c)[lo + lane];
            *x = *x + v * dt;
        }
    }
    __syncthreads();
}

__device__ void drift(const SubtreeArgs& a, int lo, int n, double dt)
{
    axpy3(a, lo, n, R_RX, R_VX, dt);
}

__device__ void kick_n(const SubtreeArgs& a, int lo, int n, double dt)
{
    axpy3(a, lo, n, R_VX, R_AX, dt);
}

// FewBody.evolve or the scheme's evolve sequence on [lo, lo + n).
__device__ void evolve(const SubtreeArgs& a, int lo, int n, double dt)
{
    if (n == 0) return;
    if (n == 1) {
        drift(a, lo, n, dt);
        return;
    }
    if (n == 2) {
        if (threadIdx.x == 0) {
            kep::Counters cnt;
            kep::advance_pair(lo, lo + 1, field(a, R_MASS), field(a, R_RX), field(a, R_RY),
                              field(a, R_RZ), field(a, R_E2), field(a, R_VX), field(a, R_VY),
                              field(a, R_VZ), dt, cnt);
            if (cnt.fail) atomicOr(a.status, 1);
        }
        __syncthreads();
        return;
    }
    for (int q = 0; q < a.nevolve; ++q) {
        const double h = a.evolve_c[q] * dt;
        if (a.evolve_op[q] == 0) {
            drift(a, lo, n, h);
        } else {
            acc_op(a, lo, n, lo, n);
            kick_n(a, lo, n, h);
        }
    }
}

// Stable partition of [lo, lo + n) into |tstep| > |tau| first; returns the
// slow count.  Every array moves with its particle.
__device__ int partition(const SubtreeArgs& a, int lo, int n, double tau)
{
    const int lane = static_cast<int>(threadIdx.x);
    const bool in = lane < n;
    const bool slow = in && fabs(field(a, R_TSTEP)[lo + (in ? lane : 0)]) > fabs(tau);
    const unsigned long long mask = __ballot(slow);
    const int ns = __popcll(mask);
    if (ns == 0 || ns == n) return ns;
    const unsigned long long below = (1ull << lane) - 1ull;
    const int rs = __popcll(mask & below);
    const int dst = slow ? rs : ns + (lane - rs);
    unsigned long long v[kSubArrays];
    if (in) {
        for (int q = 0; q < a.narr; ++q) v[q] = a.arr[q][lo + lane];
    }
    __syncthreads();
    if (in) {
        for (int q = 0; q < a.narr; ++q) a.arr[q][lo + dst] = v[q];
    }
    __syncthreads();
    return ns;
}

__global__ __launch_bounds__(kSubMax) void sia_subtree_kernel(SubtreeArgs a)
{
    __shared__ Frame stack[kSubDepth];
    const int lane = static_cast<int>(threadIdx.x);
    int sp = 0;
    int visits = 0;
    if (lane == 0) stack[0] = Frame{0, a.n, 0, -1, a.tau};
    sp = 1;
    __syncthreads();
    while (sp > 0) {
        Frame f = stack[sp - 1];
        if (f.pc < 0) {
            if (f.n == 0) {
                --sp;
                continue;
            }
            if (++visits > kSubVisits) {
                if (lane == 0) atomicOr(a.status, 3);
                return;
            }
            tstep_op(a, f.lo, f.n);
            f.ns = f.n <= 2 ? f.n : partition(a, f.lo, f.n, f.tau);
            f.pc = 0;
        }
        if (f.pc < a.nbridge) {
            const int q = f.pc++;
            const double h = a.bridge_c[q] * f.tau;
            const int nf = f.n - f.ns;
            if (lane == 0) stack[sp - 1] = f;
            __syncthreads();
            if (a.bridge_op[q] == 0) {
                evolve(a, f.lo, f.ns, h);
                if (nf > 0) {
                    if (sp == kSubDepth) {
                        if (lane == 0) atomicOr(a.status, 2);
                        return;
                    }
                    if (lane == 0) stack[sp] = Frame{f.lo + f.ns, nf, 0, -1, h};
                    ++sp;
                    __syncthreads();
                }
            } else if (f.ns > 0 && nf > 0) {
                acc_op(a, f.lo, f.ns, f.lo + f.ns, nf);
                acc_op(a, f.lo + f.ns, nf, f.lo, f.ns);
                kick_n(a, f.lo, f.ns, h);
                kick_n(a, f.lo + f.ns, nf, h);
            }
            continue;
        }
        // level done
        if (f.n - f.ns == 0 && lane == 0) *a.t_curr = *a.t_curr + f.tau;
        if (lane < f.ns) {
            const int k = f.lo + lane;
            field(a, R_TSTEP)[k] = f.tau;
            field(a, R_TIME)[k] = field(a, R_TIME)[k] + f.tau;
            long long* nstep = reinterpret_cast<long long*>(a.arr[a.role[R_NSTEP]]);
            nstep[k] = nstep[k] + 1;
        }
        __syncthreads();
        --sp;
    }
}

int sia_subtree_call(int n, void* const* arrays, int narrays, const int* roles, int acc_mode,
                     int tstep_mode, double eta, double tau, int nbridge, const int* bridge_op,
                     const double* bridge_c, int nevolve, const int* evolve_op,
                     const double* evolve_c, double* t_curr, unsigned long long* pairs,
                     int* status, void* stream)
{
    if (n < 0 || n > kSubMax) return set_error(TUPAN_ERR_ARG, "subtree size must be 0..64");
    if (narrays < R_NROLES || narrays > kSubArrays)
        return set_error(TUPAN_ERR_ARG, "bad array count");
    if (nbridge < 1 || nbridge > kSubOps || nevolve < 1 || nevolve > kSubOps)
        return set_error(TUPAN_ERR_ARG, "bad operator sequence length");
    if (!arrays || !roles || !bridge_op || !bridge_c || !evolve_op || !evolve_c || !t_curr ||
        !pairs || !status)
        return set_error(TUPAN_ERR_ARG, "null argument");
    for (int m : {acc_mode, tstep_mode})
        if (m != TUPAN_MODE_FP64 && m != TUPAN_MODE_SPLIT32)
            return set_error(TUPAN_ERR_ARG, "subtree modes are fp64 or split32");
    SubtreeArgs a{};
    a.narr = narrays;
    for (int q = 0; q < narrays; ++q) {
        if (!arrays[q]) return set_error(TUPAN_ERR_ARG, "null array");
        a.arr[q] = static_cast<unsigned long long*>(arrays[q]);
    }
    for (int r = 0; r < R_NROLES; ++r) {
        if (roles[r] < 0 || roles[r] >= narrays) return set_error(TUPAN_ERR_ARG, "bad role");
        a.role[r] = roles[r];
    }
    for (int q = 0; q < nbridge; ++q) {
        a.bridge_op[q] = bridge_op[q];
        a.bridge_c[q] = bridge_c[q];
    }
    for (int q = 0; q < nevolve; ++q) {
        a.evolve_op[q] = evolve_op[q];
        a.evolve_c[q] = evolve_c[q];
    }
    a.n = n;
    a.acc_mode = acc_mode;
    a.tstep_mode = tstep_mode;
    a.eta = eta;
    a.tau = tau;
    a.nbridge = nbridge;
    a.nevolve = nevolve;
    a.t_curr = t_curr;
    a.pairs = pairs;
    a.status = status;
    if (n == 0) return TUPAN_OK;
    hipLaunchKernelGGL(sia_subtree_kernel, dim3(1), dim3(kSubMax), 0,
                       static_cast<hipStream_t>(stream), a);
    return check_launch("sia_subtree");
}

}  // namespace tupan

// ---------------------------------------------------------------------------
// Device-driven shared-step SIA steps: the 's' (fixed step) and 'a'
// (adaptive power-of-two block step) variants of sia.py:1089-1123, in which
// every particle is slow at every step, so a step is the scheme's bridge
// sequence of evolve(all, c_b tau) calls -- each the scheme's drift/kick
// sequence, a kick being Acc(all, all) and v += a h -- followed by the
// bookkeeping.  One call enqueues a whole step with the step size in the
// device control block (block_step.h), so a sequence of calls can be
// captured as one HIP graph.  Acc goes to scratch and reaches the live
// a arrays in the kick, and Tstep ('a') to scratch and the live tstepij in
// the bookkeeping, so a step past t_end (tau = 0) changes no live array.
// Every update is the host driver's elementwise operation, one rounding at
// a time: h = c_e * (c_b * tau) as the driver forms it, then x + v * h.
// ---------------------------------------------------------------------------
namespace tupan {

namespace {

// Drifts and kicks of one shared step: the bridge's drifts times the evolve
// sequence (sia69: 10 x 19 = 190).
constexpr int kStepOps = 1024;

struct SiaStepArrays {
    double *m, *e2, *r[3], *v[3], *a[3];
    double *tstep, *tstepij, *time;
    long long* nstep;
    double *as[3], *ts, *tsij;  // scratch
};

__global__ __launch_bounds__(256) void sia_update_kernel(int n, SiaStepArrays s, double cb,
                                                         double ce, int kick,
                                                         const double* ctl)
{
    const int i = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    const double tau = ctl[kCtlTau];
    if (i >= n || tau == 0) return;
    const double h = ce * (cb * tau);
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        if (kick) {
            const double a = s.as[d][i];
            s.a[d][i] = a;
            s.v[d][i] = s.v[d][i] + a * h;
        } else {
            s.r[d][i] = s.r[d][i] + s.v[d][i] * h;
        }
    }
}

__global__ __launch_bounds__(256) void sia_end_kernel(int n, SiaStepArrays s, int adaptive,
                                                      double* ctl)
{
    const int i = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    const double tau = ctl[kCtlTau];
    if (tau == 0) return;
    if (i < n) {
        if (adaptive) s.tstepij[i] = s.tsij[i];
        s.tstep[i] = tau;
        s.time[i] = s.time[i] + tau;
        s.nstep[i] = s.nstep[i] + 1;
    }
    if (i == 0) ctl[kCtlT] = ctl[kCtlT] + tau;
}

}  // namespace

int sia_step_call(int adaptive, int acc_mode, int tstep_mode, size_t n, void* const* arrays,
                  double* scratch, double* control, double eta, int nops, const int* ops,
                  const double* coef_bridge, const double* coef_evolve, void* stream)
{
    for (int m : {acc_mode, tstep_mode})
        if (m != TUPAN_MODE_FP64 && m != TUPAN_MODE_SPLIT32 && m != TUPAN_MODE_IEEE64)
            return set_error(TUPAN_ERR_ARG, "unknown force mode");
    if (nops < 1 || nops > kStepOps)
        return set_error(TUPAN_ERR_ARG, "bad operator sequence length (1..1024)");
    if (!arrays || !scratch || !control || !ops || !coef_bridge || !coef_evolve)
        return set_error(TUPAN_ERR_ARG, "null argument");
    for (int q = 0; q < nops; ++q)
        if (ops[q] != 0 && ops[q] != 1) return set_error(TUPAN_ERR_ARG, "operator must be 0 or 1");
    if (n > static_cast<size_t>(INT32_MAX)) return set_error(TUPAN_ERR_ARG, "n exceeds 2^31-1");
    if (n < 3)
        return set_error(TUPAN_ERR_ARG, "device SIA steps need at least 3 particles "
                                        "(the reference evolves 1 or 2 by FewBody)");
    for (int q = 0; q < TUPAN_SIA_NROLES; ++q)
        if (!arrays[q]) return set_error(TUPAN_ERR_ARG, "null particle array");
    SiaStepArrays s{};
    auto d = [&](int q) { return static_cast<double*>(arrays[q]); };
    s.m = d(TUPAN_SIA_MASS);
    s.e2 = d(TUPAN_SIA_EPS2);
    for (int k = 0; k < 3; ++k) {
        s.r[k] = d(TUPAN_SIA_RX + k);
        s.v[k] = d(TUPAN_SIA_VX + k);
        s.a[k] = d(TUPAN_SIA_AX + k);
        s.as[k] = scratch + k * n;
    }
    s.tstep = d(TUPAN_SIA_TSTEP);
    s.tstepij = d(TUPAN_SIA_TSTEPIJ);
    s.time = d(TUPAN_SIA_TIME);
    s.nstep = static_cast<long long*>(arrays[TUPAN_SIA_NSTEP]);
    s.ts = scratch + 3 * n;
    s.tsij = scratch + 4 * n;
    hipStream_t st = static_cast<hipStream_t>(stream);
    const int ni = static_cast<int>(n);
    const unsigned grid = static_cast<unsigned>((n + 255) / 256);
    const double* in[8] = {s.m, s.r[0], s.r[1], s.r[2], s.e2, s.v[0], s.v[1], s.v[2]};
    if (adaptive) {
        double* out[2] = {s.ts, s.tsij};
        const int rc = pairwise_call<double>(K_TSTEP, tstep_mode, n, in, n, in, out, eta, st);
        if (rc != TUPAN_OK) return rc;
    }
    hipLaunchKernelGGL(block_tau_kernel, dim3(1), dim3(kTauThreads), 0, st, ni, s.ts, adaptive,
                       control);
    int rc = check_launch("block_tau");
    for (int q = 0; q < nops && rc == TUPAN_OK; ++q) {
        if (ops[q] == 1) {
            double* out[3] = {s.as[0], s.as[1], s.as[2]};
            rc = pairwise_call<double>(K_ACC, acc_mode, n, in, n, in, out, 0.0, st);
            if (rc != TUPAN_OK) break;
        }
        hipLaunchKernelGGL(sia_update_kernel, dim3(grid), dim3(256), 0, st, ni, s,
                           coef_bridge[q], coef_evolve[q], ops[q], control);
        rc = check_launch("sia_update");
    }
    if (rc != TUPAN_OK) return rc;
    hipLaunchKernelGGL(sia_end_kernel, dim3(grid), dim3(256), 0, st, ni, s, adaptive, control);
    return check_launch("sia_end");
}

}  // namespace tupan
