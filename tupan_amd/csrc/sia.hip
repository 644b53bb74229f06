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

// The whole recursion of one subset of at most 64 particles, run by one
// wave (the body of sia_subtree_kernel; also driven by the device-scheduled
// block steps below).
__device__ void subtree_body(const SubtreeArgs& a)
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

__global__ __launch_bounds__(kSubMax) void sia_subtree_kernel(SubtreeArgs a)
{
    subtree_body(a);
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

// ===========================================================================
// Device-scheduled hierarchical SIA steps (the 'h' variants,
// sia.py:1089-1123), with no host read of level sizes.
//
// The host driver visits each level from the host: Tstep on the level's set,
// the stable [slow | fast] partition, then the scheme's bridge -- evolve the
// slow part (drifts, and kicks with Acc(slow, slow)), recurse on the fast
// part, slow<->fast kicks -- then the bookkeeping; sets of at most 64 go to
// sia_subtree_kernel.  Here the same walk is a state machine in device
// memory (LvlCtl: a stack of level frames and a queue of pending
// operations), advanced by lvl_sched_kernel, one operation per "slot".  A
// slot is the scheduler followed by one launch of every worker kernel: the
// tile engine in each tile shape the tuning table uses for Tstep and Acc,
// the small-set kernel and both split combines, the four partition passes,
// the elementwise updates, the subtree kernel and the few-body kernel.  Each
// worker reads the descriptor the scheduler wrote and exits unless the slot
// is its own, so the launch sequence is fixed and a run of slots can be
// captured as one HIP graph and replayed until the step's stack is empty.
//
// Every operation is the host driver's, in its order, with its launch plan:
// the scheduler plans each Tstep and Acc call exactly as pairwise_call does
// (tuned shape and workgroup target per size bucket, the small-set path,
// plan_launch, the combine), and the workers run the same kernel bodies on
// the same arguments, so the device-scheduled step equals the host driver
// bit for bit.
// ===========================================================================
namespace tupan {

namespace {

constexpr int kLvlDepth = 64;        // level frames
constexpr int kLvlQueue = 128;       // pending operations of one bridge entry
constexpr int kLvlIter = 1024;       // scheduler transitions per slot
constexpr int kLvlVisits = 1 << 20;  // level visits per step
constexpr int kPartPer = 1024;       // particles per partition block (256 x 4)
constexpr int kCandSmall = kNumShapes;  // candidate id of the small-set kernel

enum { LK_NONE = 0, LK_PAIR = 1, LK_PART = 2, LK_ELEM = 3, LK_SUBTREE = 4, LK_FEW = 5 };
enum { MT_TSTEP, MT_ACC, MT_PART, MT_DRIFT, MT_KICK, MT_FEW, MT_PUSH };
enum { EL_DRIFT = 0, EL_KICK = 1, EL_BOOK = 2 };
// status bits beyond the subtree's 1..3
constexpr int kLvlErrQueue = 4, kLvlErrPlan = 8, kLvlErrIter = 16;

struct Micro {
    int type, lo, n, lo2, n2;
    double h;
};

struct LvlFrame {
    int lo, n, ns, pc;
    double tau;
};

}  // namespace

struct LvlCtl {
    double step[5];  // t_curr, t_end, eta, tau, status (block_step.h)
    // walk
    LvlFrame stack[kLvlDepth];
    Micro queue[kLvlQueue];
    int sp, qh, qt;
    int done, status, visits, ops;  // ops: slots that did work this step
    // this slot's operation
    int kind, op, cand, grid, comb, comb_grid, elem;
    int lo, n, ns;
    double h, tau;
    PairArgs<double> pa;
    // configuration (set once)
    int N, narr, acc_mode, tstep_mode, nbridge, nevolve, cu_count;
    double eta;
    unsigned long long* arr[kSubArrays];
    int role[R_NROLES];
    int bridge_op[kSubOps];
    double bridge_c[kSubOps];
    int evolve_op[kSubOps];
    double evolve_c[kSubOps];
    int tshape[2][3], twg[2][3];  // [0 Tstep, 1 Acc][size bucket]
    double* partial;
    unsigned long long* tmp;  // narr x N, partition staging
    int* blk;                 // partition block counts, then offsets
    unsigned long long pairs[2];
};

namespace {

__device__ __forceinline__ double* lvl_field(const LvlCtl* c, int role)
{
    return reinterpret_cast<double*>(c->arr[c->role[role]]);
}

__device__ __forceinline__ int lvl_bucket(int ni, int nj)
{
    const double p = static_cast<double>(ni) * static_cast<double>(nj);
    if (p < 4194304.0) return 0;
    if (p < 1073741824.0) return 1;
    return 2;
}

// The launch plan pairwise_call makes for Tstep (op 0) or Acc (op 1) of
// i in [ilo, ilo + ni) against j in [jlo, jlo + nj), written to the
// descriptor.  Returns false if the planned shape has no worker in the
// slots (the tuning table changed since the slots were enqueued).
__device__ void lvl_plan(LvlCtl* c, int op, int ilo, int ni, int jlo, int nj)
{
    PairArgs<double>& a = c->pa;
    const int fr[8] = {R_MASS, R_RX, R_RY, R_RZ, R_E2, R_VX, R_VY, R_VZ};
    for (int f = 0; f < 14; ++f) {
        a.ip[f] = nullptr;
        a.jp[f] = nullptr;
    }
    for (int f = 0; f < 8; ++f) {
        a.ip[f] = lvl_field(c, fr[f]) + ilo;
        a.jp[f] = lvl_field(c, fr[f]) + jlo;
    }
    for (int k = 0; k < 8; ++k) a.out[k] = nullptr;
    a.coincide = nullptr;
    if (op == 0) {
        a.out[0] = lvl_field(c, R_TSTEP) + ilo;
        a.out[1] = lvl_field(c, R_TSTEPIJ) + ilo;
    } else {
        a.out[0] = lvl_field(c, R_AX) + ilo;
        a.out[1] = lvl_field(c, R_AY) + ilo;
        a.out[2] = lvl_field(c, R_AZ) + ilo;
    }
    a.ni = ni;
    a.nj = nj;
    a.eta = op == 0 ? c->eta : 0.0;
    a.exact = 0;
    a.order = 0;
    for (int k = 0; k < 5; ++k) a.clight[k] = 0.0;
    a.partial = nullptr;
    c->kind = LK_PAIR;
    c->op = op;
    c->comb = 0;
    c->comb_grid = 0;
    c->pairs[op] += static_cast<unsigned long long>(ni) * static_cast<unsigned long long>(nj);
    if (nj > kSmallMinJ && nj <= kSmallMaxJ &&
        static_cast<long long>(ni) * nj <= kSmallPairs) {
        a.nsplit = 1;
        a.jchunk = nj;
        c->cand = kCandSmall;
        constexpr int ipb = kSmallBlock / kSmallGroups;
        c->grid = (ni + ipb - 1) / ipb;
        return;
    }
    const int b = lvl_bucket(ni, nj);
    const int shape = c->tshape[op][b];
    const int wg = c->twg[op][b];
    const int BLOCK = kLightShapeDims[shape][0], WPT = kLightShapeDims[shape][1],
              TILE = kLightShapeDims[shape][2];
    const int target = (wg > 0 ? wg : 4) * c->cu_count;
    const int min_chunk = min_jchunk(nj, TILE, false);
    const Launch L = plan_launch(ni, nj, BLOCK, WPT, TILE, target, 64, min_chunk);
    a.nsplit = L.nsplit;
    a.jchunk = L.jchunk;
    c->cand = shape;
    c->grid = L.nblock_i * L.nsplit;
    if (L.nsplit > 1) {
        a.partial = c->partial;
        if (ni <= kWideReduceMaxNi) {
            c->comb = 1;
            c->comb_grid = (ni * 8 + 255) / 256;
        } else {
            c->comb = 2;
            c->comb_grid = (ni + 255) / 256;
        }
    }
}

__device__ bool lvl_push_micro(LvlCtl* c, Micro m)
{
    if (c->qt - c->qh >= kLvlQueue) {
        c->status |= kLvlErrQueue;
        return false;
    }
    c->queue[c->qt % kLvlQueue] = m;
    ++c->qt;
    return true;
}

// One step of the walk: sets the descriptor of this slot's operation (or
// LK_NONE once the stack is empty or the walk failed).
__device__ void lvl_sched(LvlCtl* c)
{
    c->kind = LK_NONE;
    if (c->done) return;
    for (int it = 0; it < kLvlIter; ++it) {
        if (c->status) {
            c->done = 1;
            return;
        }
        if (c->qh < c->qt) {
            const Micro m = c->queue[c->qh % kLvlQueue];
            ++c->qh;
            switch (m.type) {
            case MT_PUSH:
                if (c->sp == kLvlDepth) {
                    c->status |= 2;
                    continue;
                }
                c->stack[c->sp] = LvlFrame{m.lo, m.n, 0, -1, m.h};
                ++c->sp;
                continue;
            case MT_TSTEP:
                lvl_plan(c, 0, m.lo, m.n, m.lo, m.n);
                return;
            case MT_ACC:
                lvl_plan(c, 1, m.lo, m.n, m.lo2, m.n2);
                return;
            case MT_PART:
                c->kind = LK_PART;
                c->lo = m.lo;
                c->n = m.n;
                c->tau = m.h;
                c->ns = 0;
                return;
            case MT_DRIFT:
            case MT_KICK:
                c->kind = LK_ELEM;
                c->elem = m.type == MT_DRIFT ? EL_DRIFT : EL_KICK;
                c->lo = m.lo;
                c->n = m.n;
                c->h = m.h;
                return;
            case MT_FEW:
                c->kind = LK_FEW;
                c->lo = m.lo;
                c->n = m.n;
                c->h = m.h;
                return;
            default:
                continue;
            }
        }
        if (c->sp == 0) {
            c->done = 1;
            return;
        }
        LvlFrame& f = c->stack[c->sp - 1];
        if (f.pc < 0) {
            if (f.n == 0) {
                --c->sp;
                continue;
            }
            if (f.n <= kSubMax) {
                c->kind = LK_SUBTREE;
                c->lo = f.lo;
                c->n = f.n;
                c->tau = f.tau;
                --c->sp;
                return;
            }
            if (++c->visits > kLvlVisits) {
                c->status |= 3;
                continue;
            }
            // Tstep of the level's set, then its [slow | fast] partition
            // (the partition writes f.ns)
            lvl_push_micro(c, Micro{MT_TSTEP, f.lo, f.n, 0, 0, 0.0});
            lvl_push_micro(c, Micro{MT_PART, f.lo, f.n, 0, 0, f.tau});
            f.pc = 0;
            continue;
        }
        if (f.pc < c->nbridge) {
            const int q = f.pc++;
            const double h = c->bridge_c[q] * f.tau;
            const int ns = f.ns, nf = f.n - f.ns;
            if (c->bridge_op[q] == 0) {
                // sf_drift: evolve(slow, h), then recurse(fast, h)
                if (ns > 0 && ns <= 2) {
                    lvl_push_micro(c, Micro{MT_FEW, f.lo, ns, 0, 0, h});
                } else if (ns > 2) {
                    for (int e = 0; e < c->nevolve; ++e) {
                        const double he = c->evolve_c[e] * h;
                        if (c->evolve_op[e] == 0) {
                            lvl_push_micro(c, Micro{MT_DRIFT, f.lo, ns, 0, 0, he});
                        } else {
                            lvl_push_micro(c, Micro{MT_ACC, f.lo, ns, f.lo, ns, 0.0});
                            lvl_push_micro(c, Micro{MT_KICK, f.lo, ns, 0, 0, he});
                        }
                    }
                }
                if (nf > 0) lvl_push_micro(c, Micro{MT_PUSH, f.lo + ns, nf, 0, 0, h});
            } else if (ns > 0 && nf > 0) {
                // sf_kick: Acc(slow <- fast), Acc(fast <- slow), both kicks
                lvl_push_micro(c, Micro{MT_ACC, f.lo, ns, f.lo + ns, nf, 0.0});
                lvl_push_micro(c, Micro{MT_ACC, f.lo + ns, nf, f.lo, ns, 0.0});
                lvl_push_micro(c, Micro{MT_KICK, f.lo, ns, 0, 0, h});
                lvl_push_micro(c, Micro{MT_KICK, f.lo + ns, nf, 0, 0, h});
            }
            continue;
        }
        // the level is done: time and the slow part's bookkeeping
        if (f.n - f.ns == 0) c->step[kCtlT] = c->step[kCtlT] + f.tau;
        c->kind = LK_ELEM;
        c->elem = EL_BOOK;
        c->lo = f.lo;
        c->n = f.ns;
        c->tau = f.tau;
        --c->sp;
        return;
    }
    c->status |= kLvlErrIter;
    c->done = 1;
}

__global__ __launch_bounds__(64) void lvl_sched_kernel(LvlCtl* c)
{
    if (threadIdx.x != 0) return;
    lvl_sched(c);
    if (c->kind != LK_NONE) ++c->ops;
}

// Start of a step: the base step is in step[kCtlTau] (block_tau_kernel with
// no adaptive part, get_base_tstep); the walk starts at the whole set.
__global__ __launch_bounds__(64) void lvl_begin_kernel(LvlCtl* c)
{
    if (threadIdx.x != 0) return;
    c->sp = 0;
    c->qh = c->qt = 0;
    c->visits = 0;
    c->ops = 0;
    c->kind = LK_NONE;
    const double tau = c->step[kCtlTau];
    if (c->status == 0 && c->step[kCtlStatus] == 0 && tau != 0) {
        c->stack[0] = LvlFrame{0, c->N, 0, -1, tau};
        c->sp = 1;
        c->done = 0;
    } else {
        if (c->step[kCtlStatus] != 0) c->status |= 32;
        c->done = 1;
    }
}

template <class Op, int S>
__global__ __launch_bounds__(kLightShapeDims[S][0]) void lvl_pair_kernel(const LvlCtl* c, int op)
{
    if (c->kind != LK_PAIR || c->op != op || c->cand != S) return;
    const int bid = static_cast<int>(blockIdx.x);
    if (bid >= c->grid) return;
    pair_body<Op, kLightShapeDims[S][0], kLightShapeDims[S][1], kLightShapeDims[S][2],
              kLightShapeDims[S][3]>(c->pa, bid);
}

template <class Op>
__global__ __launch_bounds__(kSmallBlock) void lvl_small_kernel(const LvlCtl* c, int op)
{
    if (c->kind != LK_PAIR || c->op != op || c->cand != kCandSmall) return;
    const int bid = static_cast<int>(blockIdx.x);
    if (bid >= c->grid) return;
    pair_small_body<Op, kSmallBlock, kSmallGroups, small_tile<Op>()>(c->pa, bid);
}

template <class Op>
__global__ __launch_bounds__(256) void lvl_combine_kernel(const LvlCtl* c, int op, int wide)
{
    if (c->kind != LK_PAIR || c->op != op || c->comb != (wide ? 1 : 2)) return;
    const int bid = static_cast<int>(blockIdx.x);
    if (bid >= c->comb_grid) return;
    if (wide) {
        reduce_wide_body<Op, 8>(c->pa, bid);
    } else {
        reduce_body<Op>(c->pa, bid);
    }
}

// Partition of [lo, lo + n) into |tstep| > |tau| first, stable, every array
// moving with its particle: per-block slow counts, their exclusive scan (and
// the slow count into the level's frame), the scatter into staging, and the
// copy back.  A level with no slow or no fast particle keeps its order, as
// the host's partition does.
__device__ __forceinline__ bool lvl_is_slow(const LvlCtl* c, int k)
{
    return fabs(lvl_field(c, R_TSTEP)[k]) > fabs(c->tau);
}

__global__ __launch_bounds__(256) void lvl_part_count_kernel(LvlCtl* c)
{
    __shared__ int red[256];
    if (c->kind != LK_PART) return;
    const int nb = (c->n + kPartPer - 1) / kPartPer;
    const int b = static_cast<int>(blockIdx.x);
    if (b >= nb) return;
    const int t = static_cast<int>(threadIdx.x);
    int cnt = 0;
    for (int u = 0; u < 4; ++u) {
        const int i = b * kPartPer + t * 4 + u;
        if (i < c->n && lvl_is_slow(c, c->lo + i)) ++cnt;
    }
    red[t] = cnt;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (t < w) red[t] += red[t + w];
        __syncthreads();
    }
    if (t == 0) c->blk[b] = red[0];
}

__global__ __launch_bounds__(1024) void lvl_part_scan_kernel(LvlCtl* c)
{
    __shared__ int sh[1024];
    __shared__ int carry;
    if (c->kind != LK_PART) return;
    const int nb = (c->n + kPartPer - 1) / kPartPer;
    const int t = static_cast<int>(threadIdx.x);
    if (t == 0) carry = 0;
    __syncthreads();
    for (int base = 0; base < nb; base += 1024) {
        const int v = base + t < nb ? c->blk[base + t] : 0;
        sh[t] = v;
        __syncthreads();
        for (int off = 1; off < 1024; off <<= 1) {
            const int x = t >= off ? sh[t - off] : 0;
            __syncthreads();
            sh[t] += x;
            __syncthreads();
        }
        if (base + t < nb) c->blk[base + t] = carry + sh[t] - v;  // exclusive
        __syncthreads();
        if (t == 1023) carry += sh[1023];
        __syncthreads();
    }
    if (t == 0) {
        c->ns = carry;
        c->stack[c->sp - 1].ns = carry;
    }
}

__global__ __launch_bounds__(256) void lvl_part_scatter_kernel(LvlCtl* c)
{
    __shared__ int sh[256];
    if (c->kind != LK_PART) return;
    const int n = c->n, ns = c->ns;
    if (ns == 0 || ns == n) return;
    const int nb = (n + kPartPer - 1) / kPartPer;
    const int b = static_cast<int>(blockIdx.x);
    if (b >= nb) return;
    const int t = static_cast<int>(threadIdx.x);
    bool sl[4];
    int cnt = 0;
    for (int u = 0; u < 4; ++u) {
        const int i = b * kPartPer + t * 4 + u;
        sl[u] = i < n && lvl_is_slow(c, c->lo + i);
        cnt += sl[u] ? 1 : 0;
    }
    // inclusive scan of the per-thread counts over the block
    sh[t] = cnt;
    __syncthreads();
    for (int off = 1; off < 256; off <<= 1) {
        const int x = t >= off ? sh[t - off] : 0;
        __syncthreads();
        sh[t] += x;
        __syncthreads();
    }
    int slow_before = c->blk[b] + sh[t] - cnt;
    const size_t N = static_cast<size_t>(c->N);
    for (int u = 0; u < 4; ++u) {
        const int i = b * kPartPer + t * 4 + u;
        if (i >= n) break;
        const int dst = sl[u] ? slow_before : ns + (i - slow_before);
        if (sl[u]) ++slow_before;
        for (int q = 0; q < c->narr; ++q) c->tmp[q * N + dst] = c->arr[q][c->lo + i];
    }
}

__global__ __launch_bounds__(256) void lvl_part_copy_kernel(LvlCtl* c)
{
    if (c->kind != LK_PART) return;
    const int n = c->n, ns = c->ns;
    if (ns == 0 || ns == n) return;
    const int i = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= n) return;
    const size_t N = static_cast<size_t>(c->N);
    for (int q = 0; q < c->narr; ++q) c->arr[q][c->lo + i] = c->tmp[q * N + i];
}

// Drift r += v h, kick v += a h, or the bookkeeping of a level's slow part
// (tstep = tau, time += tau, nstep += 1), as the host's elementwise updates.
__global__ __launch_bounds__(256) void lvl_elem_kernel(LvlCtl* c)
{
    if (c->kind != LK_ELEM) return;
    const int i = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= c->n) return;
    const int k = c->lo + i;
    if (c->elem == EL_BOOK) {
        lvl_field(c, R_TSTEP)[k] = c->tau;
        lvl_field(c, R_TIME)[k] = lvl_field(c, R_TIME)[k] + c->tau;
        long long* nstep = reinterpret_cast<long long*>(c->arr[c->role[R_NSTEP]]);
        nstep[k] = nstep[k] + 1;
        return;
    }
    const int xr = c->elem == EL_DRIFT ? R_RX : R_VX;
    const int vr = c->elem == EL_DRIFT ? R_VX : R_AX;
    const double h = c->h;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        double* x = lvl_field(c, xr + d) + k;
        *x = *x + lvl_field(c, vr + d)[k] * h;
    }
}

// A set of at most 64 particles: its whole recursion in one wave.
__global__ __launch_bounds__(kSubMax) void lvl_subtree_kernel(LvlCtl* c)
{
    if (c->kind != LK_SUBTREE) return;
    SubtreeArgs a{};
    a.narr = c->narr;
    for (int q = 0; q < c->narr; ++q) a.arr[q] = c->arr[q] + c->lo;
    for (int r = 0; r < R_NROLES; ++r) a.role[r] = c->role[r];
    a.n = c->n;
    a.acc_mode = c->acc_mode;
    a.tstep_mode = c->tstep_mode;
    a.eta = c->eta;
    a.tau = c->tau;
    a.nbridge = c->nbridge;
    a.nevolve = c->nevolve;
    for (int q = 0; q < c->nbridge; ++q) {
        a.bridge_op[q] = c->bridge_op[q];
        a.bridge_c[q] = c->bridge_c[q];
    }
    for (int q = 0; q < c->nevolve; ++q) {
        a.evolve_op[q] = c->evolve_op[q];
        a.evolve_c[q] = c->evolve_c[q];
    }
    a.t_curr = &c->step[kCtlT];
    a.pairs = c->pairs;
    a.status = &c->status;
    subtree_body(a);
}

// evolve of a slow part of one or two particles (FewBody.evolve): a drift,
// or the exact two-body advance.
__global__ __launch_bounds__(64) void lvl_few_kernel(LvlCtl* c)
{
    if (c->kind != LK_FEW || threadIdx.x != 0) return;
    const int lo = c->lo;
    if (c->n == 1) {
        for (int d = 0; d < 3; ++d) {
            double* x = lvl_field(c, R_RX + d) + lo;
            *x = *x + lvl_field(c, R_VX + d)[lo] * c->h;
        }
        return;
    }
    kep::Counters cnt;
    kep::advance_pair(lo, lo + 1, lvl_field(c, R_MASS), lvl_field(c, R_RX), lvl_field(c, R_RY),
                      lvl_field(c, R_RZ), lvl_field(c, R_E2), lvl_field(c, R_VX),
                      lvl_field(c, R_VY), lvl_field(c, R_VZ), c->h, cnt);
    if (cnt.fail) atomicOr(&c->status, 1);
}

}  // namespace

// Host side: one plan per particle system and method.
struct LevelsPlan {
    LvlCtl* ctl = nullptr;  // device
    double* partial = nullptr;
    unsigned long long* tmp = nullptr;
    int* blk = nullptr;
    int N = 0;
    int tstep_mode = 0, acc_mode = 0;
    int shapes[2][kNumShapes] = {};   // shapes the tuning table uses, per op
    int grid_max[2][kNumShapes] = {};
};

int sia_levels_create(int n, void* const* arrays, int narrays, const int* roles, int acc_mode,
                      int tstep_mode, double eta, int nbridge, const int* bridge_op,
                      const double* bridge_c, int nevolve, const int* evolve_op,
                      const double* evolve_c, void** plan_out)
{
    if (!plan_out) return set_error(TUPAN_ERR_ARG, "null plan pointer");
    *plan_out = nullptr;
    if (n <= kSubMax) return set_error(TUPAN_ERR_ARG, "device levels need more than 64 particles");
    if (narrays < R_NROLES || narrays > kSubArrays)
        return set_error(TUPAN_ERR_ARG, "bad array count");
    if (nbridge < 1 || nbridge > kSubOps || nevolve < 1 || nevolve > kSubOps)
        return set_error(TUPAN_ERR_ARG, "bad operator sequence length");
    if (!arrays || !roles || !bridge_op || !bridge_c || !evolve_op || !evolve_c)
        return set_error(TUPAN_ERR_ARG, "null argument");
    for (int m : {acc_mode, tstep_mode})
        if (m != TUPAN_MODE_FP64 && m != TUPAN_MODE_SPLIT32)
            return set_error(TUPAN_ERR_ARG, "device level modes are fp64 or split32");
    LvlCtl h{};
    h.N = n;
    h.narr = narrays;
    for (int q = 0; q < narrays; ++q) {
        if (!arrays[q]) return set_error(TUPAN_ERR_ARG, "null array");
        h.arr[q] = static_cast<unsigned long long*>(arrays[q]);
    }
    for (int r = 0; r < R_NROLES; ++r) {
        if (roles[r] < 0 || roles[r] >= narrays) return set_error(TUPAN_ERR_ARG, "bad role");
        h.role[r] = roles[r];
    }
    for (int q = 0; q < nbridge; ++q) {
        h.bridge_op[q] = bridge_op[q];
        h.bridge_c[q] = bridge_c[q];
    }
    for (int q = 0; q < nevolve; ++q) {
        h.evolve_op[q] = evolve_op[q];
        h.evolve_c[q] = evolve_c[q];
    }
    h.nbridge = nbridge;
    h.nevolve = nevolve;
    h.acc_mode = acc_mode;
    h.tstep_mode = tstep_mode;
    h.eta = eta;
    h.cu_count = current_device().cu_count;
    LevelsPlan* p = new LevelsPlan();
    p->N = n;
    p->tstep_mode = tstep_mode;
    p->acc_mode = acc_mode;
    // the tuning table's shapes for Tstep and Acc in these modes, and each
    // shape's largest grid over calls of at most n x n pairs
    int wg_max[2][kNumShapes] = {};
    for (int op = 0; op < 2; ++op) {
        const int kernel = op == 0 ? K_TSTEP : K_ACC;
        const int mode = op == 0 ? tstep_mode : acc_mode;
        const int slot = mode == TUPAN_MODE_SPLIT32 ? 1 : 0;
        for (int b = 0; b < 3; ++b) {
            int shape = 0, wg = 4;
            tune_get(kernel, slot, b, &shape, &wg);
            if (shape < 0 || shape >= kNumShapes) shape = 0;
            h.tshape[op][b] = shape;
            h.twg[op][b] = wg;
            p->shapes[op][shape] = 1;
            const int w = wg > 0 ? wg : 4;
            if (w > wg_max[op][shape]) wg_max[op][shape] = w;
        }
    }
    size_t part_doubles = 0;
    for (int op = 0; op < 2; ++op) {
        const size_t nacc = op == 0 ? 2 : 3;
        for (int s = 0; s < kNumShapes; ++s) {
            if (!p->shapes[op][s]) continue;
            const int ipb = kLightShapeDims[s][0] * kLightShapeDims[s][1];
            const long long target = static_cast<long long>(wg_max[op][s]) * h.cu_count;
            const long long nblk = (n + ipb - 1) / ipb;
            p->grid_max[op][s] = static_cast<int>(target + nblk + 1);
            // nsplit * ni <= min(64 ni, (target / nblock_i + 1) * nblock_i * ipb)
            const size_t a = static_cast<size_t>(64) * n;
            const size_t b = static_cast<size_t>(target + nblk) * ipb;
            const size_t need = nacc * (a < b ? a : b);
            if (need > part_doubles) part_doubles = need;
        }
    }
    const size_t nb = (static_cast<size_t>(n) + kPartPer - 1) / kPartPer;
    int rc = check_hip(hipMalloc(&p->ctl, sizeof(LvlCtl)), "levels control");
    if (rc == TUPAN_OK)
        rc = check_hip(hipMalloc(&p->partial, sizeof(double) * (part_doubles + 1)), "levels partials");
    if (rc == TUPAN_OK)
        rc = check_hip(hipMalloc(&p->tmp, sizeof(unsigned long long) * narrays * n),
                       "levels staging");
    if (rc == TUPAN_OK) rc = check_hip(hipMalloc(&p->blk, sizeof(int) * (nb + 1)), "levels blocks");
    if (rc == TUPAN_OK) {
        h.partial = p->partial;
        h.tmp = p->tmp;
        h.blk = p->blk;
        h.done = 1;
        rc = check_hip(hipMemcpy(p->ctl, &h, sizeof(LvlCtl), hipMemcpyHostToDevice), "levels setup");
    }
    if (rc != TUPAN_OK) {
        (void)hipFree(p->ctl);
        (void)hipFree(p->partial);
        (void)hipFree(p->tmp);
        (void)hipFree(p->blk);
        delete p;
        return rc;
    }
    *plan_out = p;
    return TUPAN_OK;
}

int sia_levels_destroy(void* plan)
{
    LevelsPlan* p = static_cast<LevelsPlan*>(plan);
    if (!p) return TUPAN_OK;
    const int rc = check_hip(hipDeviceSynchronize(), "levels destroy");
    (void)hipFree(p->ctl);
    (void)hipFree(p->partial);
    (void)hipFree(p->tmp);
    (void)hipFree(p->blk);
    delete p;
    return rc;
}

namespace {

template <class Op>
int lvl_enqueue_pair_workers(const LevelsPlan* p, int op, hipStream_t s)
{
    // one launch per shape the tuning table uses, the small-set kernel and
    // the two combines
#define TUPAN_LVL_SHAPE(S)                                                                   \
    if (p->shapes[op][S])                                                                    \
        hipLaunchKernelGGL((lvl_pair_kernel<Op, S>), dim3(p->grid_max[op][S]),               \
                           dim3(kLightShapeDims[S][0]), 0, s, p->ctl, op);
    TUPAN_LVL_SHAPE(0)
    TUPAN_LVL_SHAPE(1)
    TUPAN_LVL_SHAPE(2)
    TUPAN_LVL_SHAPE(3)
    TUPAN_LVL_SHAPE(4)
    TUPAN_LVL_SHAPE(5)
#undef TUPAN_LVL_SHAPE
    constexpr int ipb = kSmallBlock / kSmallGroups;
    // the small-set path has nj > kSmallMinJ, so ni <= kSmallPairs / kSmallMinJ
    const int small_grid = static_cast<int>((kSmallPairs / kSmallMinJ + ipb - 1) / ipb);
    hipLaunchKernelGGL((lvl_small_kernel<Op>), dim3(small_grid), dim3(kSmallBlock), 0, s, p->ctl,
                       op);
    hipLaunchKernelGGL((lvl_combine_kernel<Op>), dim3((kWideReduceMaxNi * 8 + 255) / 256),
                       dim3(256), 0, s, p->ctl, op, 1);
    hipLaunchKernelGGL((lvl_combine_kernel<Op>), dim3((p->N + 255) / 256), dim3(256), 0, s,
                       p->ctl, op, 0);
    return check_launch("levels pair workers");
}

}  // namespace

int sia_levels_enqueue(void* plan, int begin, int nslots, void* stream)
{
    LevelsPlan* p = static_cast<LevelsPlan*>(plan);
    if (!p) return set_error(TUPAN_ERR_ARG, "null plan");
    if (nslots < 0) return set_error(TUPAN_ERR_ARG, "negative slot count");
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (begin) {
        hipLaunchKernelGGL(block_tau_kernel, dim3(1), dim3(kTauThreads), 0, s, 0,
                           static_cast<const double*>(nullptr), 0,
                           reinterpret_cast<double*>(p->ctl));
        hipLaunchKernelGGL(lvl_begin_kernel, dim3(1), dim3(64), 0, s, p->ctl);
    }
    const int nb = (p->N + kPartPer - 1) / kPartPer;
    const int ng = (p->N + 255) / 256;
    for (int k = 0; k < nslots; ++k) {
        hipLaunchKernelGGL(lvl_sched_kernel, dim3(1), dim3(64), 0, s, p->ctl);
        int rc = p->tstep_mode == TUPAN_MODE_SPLIT32
                     ? lvl_enqueue_pair_workers<OpSplitTstep>(p, 0, s)
                     : lvl_enqueue_pair_workers<OpTstep<Fast64>>(p, 0, s);
        if (rc == TUPAN_OK)
            rc = p->acc_mode == TUPAN_MODE_SPLIT32
                     ? lvl_enqueue_pair_workers<OpSplitAcc>(p, 1, s)
                     : lvl_enqueue_pair_workers<OpAcc<Fast64>>(p, 1, s);
        if (rc != TUPAN_OK) return rc;
        hipLaunchKernelGGL(lvl_part_count_kernel, dim3(nb), dim3(256), 0, s, p->ctl);
        hipLaunchKernelGGL(lvl_part_scan_kernel, dim3(1), dim3(1024), 0, s, p->ctl);
        hipLaunchKernelGGL(lvl_part_scatter_kernel, dim3(nb), dim3(256), 0, s, p->ctl);
        hipLaunchKernelGGL(lvl_part_copy_kernel, dim3(ng), dim3(256), 0, s, p->ctl);
        hipLaunchKernelGGL(lvl_elem_kernel, dim3(ng), dim3(256), 0, s, p->ctl);
        hipLaunchKernelGGL(lvl_subtree_kernel, dim3(1), dim3(kSubMax), 0, s, p->ctl);
        hipLaunchKernelGGL(lvl_few_kernel, dim3(1), dim3(64), 0, s, p->ctl);
        rc = check_launch("levels slot");
        if (rc != TUPAN_OK) return rc;
    }
    return TUPAN_OK;
}

int sia_levels_set(void* plan, double t_curr, double t_end, double eta, int clear_pairs)
{
    LevelsPlan* p = static_cast<LevelsPlan*>(plan);
    if (!p) return set_error(TUPAN_ERR_ARG, "null plan");
    LvlCtl* d = p->ctl;
    const double step[5] = {t_curr, t_end, eta, 0.0, 0.0};
    const int state[4] = {1, 0, 0, 0};  // done, status, visits, ops
    int rc = check_hip(hipMemcpy(&d->step[0], step, sizeof(step), hipMemcpyHostToDevice),
                       "levels time");
    if (rc == TUPAN_OK)
        rc = check_hip(hipMemcpy(&d->done, state, sizeof(state), hipMemcpyHostToDevice),
                       "levels state");
    if (rc == TUPAN_OK && clear_pairs) {
        const unsigned long long z[2] = {0, 0};
        rc = check_hip(hipMemcpy(&d->pairs[0], z, sizeof(z), hipMemcpyHostToDevice),
                       "levels counters");
    }
    return rc;
}

int sia_levels_state(void* plan, double* step, int* state, unsigned long long* pairs)
{
    LevelsPlan* p = static_cast<LevelsPlan*>(plan);
    if (!p) return set_error(TUPAN_ERR_ARG, "null plan");
    LvlCtl* d = p->ctl;
    int rc = TUPAN_OK;
    if (step)
        rc = check_hip(hipMemcpy(step, &d->step[0], 5 * sizeof(double), hipMemcpyDeviceToHost),
                       "levels time");
    if (rc == TUPAN_OK && state)
        rc = check_hip(hipMemcpy(state, &d->done, 4 * sizeof(int), hipMemcpyDeviceToHost),
                       "levels state");
    if (rc == TUPAN_OK && pairs)
        rc = check_hip(hipMemcpy(pairs, &d->pairs[0], 2 * sizeof(unsigned long long),
                                 hipMemcpyDeviceToHost),
                       "levels counters");
    return rc;
}

}  // namespace tupan
