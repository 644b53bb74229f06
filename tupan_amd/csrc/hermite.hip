// Device-side Hermite steps (tupan/integrator/hermite.py:20-410 and the
// shared block step of tupan/integrator/__init__.py:48-78).
//
// One call advances the whole system by one P(EC)^2 step with every decision
// taken on the device: the step size is read from and written to a small
// control block in device memory, so a sequence of calls can be captured as
// one HIP graph and replayed with no host round trip.  Per step:
//
//   1. 'a' methods: Tstep (eta) into scratch arrays;
//   2. block_tau_kernel: the base step min(|t_end| - |t|, |eta|) and, for
//      'a' methods, the power-of-two block step below min |tstep| aligned
//      with t (get_min_block_tstep); tau = 0 once t has reached t_end, which
//      turns the rest of the step into a no-op, and also once the control
//      block's status is nonzero (1: non-finite t, t_end or eta; 2: no
//      positive finite Tstep; 3: no aligned block step), which it keeps;
//   3. AccJerk (+ SnapCrackle for orders 6, 8) of the current state into the
//      scratch a0, j0 (s0, c0); the predictor saves r0, v0 and predicts r, v;
//   4. twice: AccJerk (+ SnapCrackle) of the live state into the scratch a1,
//      j1 (s1, c1), then the corrector, which also copies them into the live
//      a, j (s, c);
//   5. hermite_end_kernel: t += tau, tstep = tau, time += tau, nstep += 1,
//      tstepij from the scratch Tstep output.
// Force evaluations write only scratch arrays and every elementwise kernel
// does nothing when tau = 0, so a step after t_end changes no live array.
//
// Order 4 on sets the small-set kernel takes (257 to 2048 particles, fp64
// or split32) fuses steps 3-5 into the three force launches: each launch's
// per-particle store runs that particle's predictor, corrector or corrector
// and bookkeeping.  Those epilogues cannot write the positions the same
// launch reads as j-particles, so the predicted state goes to the r0, v0
// scratch, the once-corrected state to the s0, c0 scratch (unused at order
// 4), and the live r, v keep the start-of-step state until the last
// corrector replaces them: five launches per step instead of nine, with the
// same operations in the same order, so the same results bit for bit.
//
// The predictor and corrector evaluate the reference's polynomials operation
// by operation as the host driver does on its arrays (one rounding per
// operation, IEEE division, no contraction: this file is compiled with
// -ffp-contract=off), so the device-driven trajectory equals the
// host-driven one bit for bit.  The step-size logic is the host's in double
// precision; its log2 is the device library's, which can differ from numpy's
// in the last bit for a minimum time step within an ulp or two of a power of
// two.
#include "block_step.h"
#include "ops.h"

namespace tupan {

namespace {

struct HermiteArrays {
    double *m, *e2, *r[3], *v[3], *a[3], *j[3], *s[3], *c[3];
    double *tstep, *tstepij, *time;
    long long* nstep;
    // scratch
    double *r0[3], *v0[3], *a0[3], *j0[3], *s0[3], *c0[3], *ts, *tsij;
    double *a1[3], *j1[3], *s1[3], *c1[3];
};

// hermite.py:20-410, predictor of H2/H4/H6/H8 for component d of particle
// i: the start-of-step r, v (live arrays) and the start forces a0, j0
// (passed in) and s0, c0, the predicted r, v to ro, vo and, when rs is
// given, the start r, v to rs, vs.
template <int Order>
__device__ __forceinline__ void predict_dim(const HermiteArrays& h, int i, int d, double tau,
                                            double a0, double j0, double* const* rs,
                                            double* const* vs, double* const* ro,
                                            double* const* vo)
{
    const double r0 = h.r[d][i], v0 = h.v[d][i];
    if (rs) {
        rs[d][i] = r0;
        vs[d][i] = v0;
    }
    double dr, dv;
    if constexpr (Order == 2) {
        dr = (a0 * tau / 2 + v0) * tau;
        dv = a0 * tau;
    } else if constexpr (Order == 4) {
        dr = ((j0 * tau / 3 + a0) * tau / 2 + v0) * tau;
        dv = (j0 * tau / 2 + a0) * tau;
    } else if constexpr (Order == 6) {
        const double s0 = h.s0[d][i];
        dr = (((s0 * tau / 4 + j0) * tau / 3 + a0) * tau / 2 + v0) * tau;
        dv = ((s0 * tau / 3 + j0) * tau / 2 + a0) * tau;
    } else {
        const double s0 = h.s0[d][i], c0 = h.c0[d][i];
        dr = ((((c0 * tau / 5 + s0) * tau / 4 + j0) * tau / 3 + a0) * tau / 2 + v0) * tau;
        dv = (((c0 * tau / 4 + s0) * tau / 3 + j0) * tau / 2 + a0) * tau;
    }
    ro[d][i] = r0 + dr;
    vo[d][i] = v0 + dv;
}

template <int Order>
__device__ __forceinline__ void predict_one(const HermiteArrays& h, int i, double tau,
                                            double* const* rs, double* const* vs,
                                            double* const* ro, double* const* vo)
{
#pragma unroll
    for (int d = 0; d < 3; ++d)
        predict_dim<Order>(h, i, d, tau, h.a0[d][i], Order >= 4 ? h.j0[d][i] : 0.0, rs, vs,
                           ro, vo);
}

template <int Order>
__global__ __launch_bounds__(256) void hermite_predict_kernel(int n, HermiteArrays h,
                                                              const double* ctl)
{
    const int i = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    const double tau = ctl[kCtlTau];
    if (i >= n || tau == 0) return;
    predict_one<Order>(h, i, tau, h.r0, h.v0, h.r, h.v);
}

// hermite.py:20-410, corrector of H2/H4/H6/H8 for component d of particle
// i (v first, then r with the new v): the start-of-step r, v from r0s, v0s,
// the start forces a0.. and the new forces a1 (j1 for order 4 passed in,
// higher derivatives read from s1, c1), which it also copies into the live
// a (j, s, c); the corrected r, v go to ro, vo.
template <int Order>
__device__ __forceinline__ void correct_dim(const HermiteArrays& h, int i, int d, double tau,
                                            double a1, double j1, double* const* r0s,
                                            double* const* v0s, double* const* ro,
                                            double* const* vo)
{
    const double r0 = r0s[d][i], v0 = v0s[d][i];
    const double a0 = h.a0[d][i];
    h.a[d][i] = a1;
    if constexpr (Order >= 4) h.j[d][i] = j1;
    if constexpr (Order >= 6) {
        h.s[d][i] = h.s1[d][i];
        h.c[d][i] = h.c1[d][i];
    }
    double v1, r1;
    if constexpr (Order == 2) {
        v1 = (a0 + a1) * tau / 2 + v0;
        r1 = (v0 + v1) * tau / 2 + r0;
    } else if constexpr (Order == 4) {
        const double j0 = h.j0[d][i];
        v1 = ((j0 - j1) * tau / 6 + (a0 + a1)) * tau / 2 + v0;
        r1 = ((a0 - a1) * tau / 6 + (v0 + v1)) * tau / 2 + r0;
    } else if constexpr (Order == 6) {
        const double j0 = h.j0[d][i];
        const double s0 = h.s0[d][i], s1 = h.s1[d][i];
        v1 = (((s0 + s1) * tau / 12 + (j0 - j1)) * tau / 5 + (a0 + a1)) * tau / 2 + v0;
        r1 = (((j0 + j1) * tau / 12 + (a0 - a1)) * tau / 5 + (v0 + v1)) * tau / 2 + r0;
    } else {
        const double j0 = h.j0[d][i];
        const double s0 = h.s0[d][i], s1 = h.s1[d][i];
        const double c0 = h.c0[d][i], c1 = h.c1[d][i];
        v1 = ((((c0 - c1) * tau / 20 + (s0 + s1)) * tau / 3 + 3 * (j0 - j1)) * tau / 14 +
              (a0 + a1)) * tau / 2 + v0;
        r1 = ((((s0 - s1) * tau / 20 + (j0 + j1)) * tau / 3 + 3 * (a0 - a1)) * tau / 14 +
              (v0 + v1)) * tau / 2 + r0;
    }
    vo[d][i] = v1;
    ro[d][i] = r1;
}

template <int Order>
__device__ __forceinline__ void correct_one(const HermiteArrays& h, int i, double tau,
                                            double* const* r0s, double* const* v0s,
                                            double* const* ro, double* const* vo)
{
#pragma unroll
    for (int d = 0; d < 3; ++d)
        correct_dim<Order>(h, i, d, tau, h.a1[d][i], Order >= 4 ? h.j1[d][i] : 0.0, r0s, v0s,
                           ro, vo);
}

template <int Order>
__global__ __launch_bounds__(256) void hermite_correct_kernel(int n, HermiteArrays h,
                                                              const double* ctl)
{
    const int i = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    const double tau = ctl[kCtlTau];
    if (i >= n || tau == 0) return;
    correct_one<Order>(h, i, tau, h.r0, h.v0, h.r, h.v);
}

// hermite.py:390-410 (do_step's bookkeeping) for particle i.
__device__ __forceinline__ void end_one(const HermiteArrays& h, int i, double tau, int adaptive)
{
    h.tstep[i] = tau;
    h.time[i] = h.time[i] + tau;
    h.nstep[i] = h.nstep[i] + 1;
    if (adaptive) h.tstepij[i] = h.tsij[i];
}

// Thread 0 of block 0 also advances t; every thread reads tau before that
// write is visible to it or not at all, since tau is not changed here.
__global__ __launch_bounds__(256) void hermite_end_kernel(int n, HermiteArrays h, int adaptive,
                                                          double* ctl)
{
    const int i = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    const double tau = ctl[kCtlTau];
    if (tau == 0) return;
    if (i < n) end_one(h, i, tau, adaptive);
    if (i == 0) ctl[kCtlT] = ctl[kCtlT] + tau;
}

// AccJerk of the small-set kernel with a Hermite-4 epilogue in its
// per-particle store, spread over the particle's lane group: lane d < 3
// runs component d, lane 3 the bookkeeping.  Stage 0: forces of the start
// state into a0, j0, then the predictor into the r0, v0 scratch; stage 1:
// forces of the predicted state into a1, j1, then the corrector into the
// s0, c0 scratch; stage 2: forces of that state into a1, j1, then the
// corrector into the live r, v and the bookkeeping.  The new forces come
// from the lanes' registers (the values the store writes), never from the
// arrays lane 0 is writing.
template <class Base, int Stage>
struct OpHermite4 : Base {
    struct Args : Base::Args {
        HermiteArrays h;
        double* ctl;
        int adaptive;
    };
    __device__ static void store_group(const Args& a, int i, const double* t, int lane)
    {
        if (lane == 0) Base::store(a, i, t);
        const double tau = a.ctl[kCtlTau];
        if (tau == 0) return;
        const HermiteArrays& h = a.h;
        if (lane < 3) {
            const int d = lane;
            const double f = d == 0 ? t[0] : (d == 1 ? t[1] : t[2]);
            const double g = d == 0 ? t[3] : (d == 1 ? t[4] : t[5]);
            if constexpr (Stage == 0) {
                predict_dim<4>(h, i, d, tau, f, g, nullptr, nullptr, h.r0, h.v0);
            } else if constexpr (Stage == 1) {
                correct_dim<4>(h, i, d, tau, f, g, h.r, h.v, h.s0, h.c0);
            } else {
                correct_dim<4>(h, i, d, tau, f, g, h.r, h.v, h.r, h.v);
            }
        } else if (lane == 3 && Stage == 2) {
            end_one(h, i, tau, a.adaptive);
            if (i == 0) a.ctl[kCtlT] = a.ctl[kCtlT] + tau;
        }
    }
};

// Tstep and AccJerk of one state in one pass of the small-set kernel: the
// two interactions read the same slot (m rx ry rz e2 vx vy vz; the split
// Tstep, which holds doubled masses, gets a copy with 2 m, exact), and each
// keeps its own sums, flushed at the same points, so both results equal
// those of the two separate calls bit for bit (the same pairs per lane in
// the same order, the same combine tree, and, in split32, the FP64
// evaluation of a lane's sums only for the interaction whose sums are not
// finite).  out[0..5] take a, j and out[6], out[7] the two Tstep outputs.
template <class AJ, class TS>
struct OpAccJerkTstep {
    using Args = PairArgs<double>;
    using Frame = typename AJ::Frame;
    using Slot = typename AJ::Slot;
    static constexpr bool kSplit = std::is_same<AJ, OpSplitAccJerk>::value;
    static_assert(std::is_same<Slot, typename TS::Slot>::value,
                  "one slot layout serves both interactions");
    struct IReg { typename AJ::IReg a; typename TS::IReg t; };
    struct Acc { typename AJ::Acc a; typename TS::Acc t; };
    static constexpr int kNacc = AJ::kNacc + TS::kNacc;
    static constexpr int kFlush = AJ::kFlush < TS::kFlush ? AJ::kFlush : TS::kFlush;
    __device__ static Frame frame(const Args& a, int ibase) { return AJ::frame(a, ibase); }
    __device__ static void load_i(const Args& a, const Frame& f, int i, IReg& r)
    {
        AJ::load_i(a, f, i, r.a);
        TS::load_i(a, f, i, r.t);
    }
    __device__ static void load_j(const Args& a, const Frame& f, int j, Slot& s)
    {
        AJ::load_j(a, f, j, s);
    }
    __device__ static void zero(Acc& c)
    {
        AJ::zero(c.a);
        TS::zero(c.t);
    }
    __device__ static __forceinline__ void interact(const IReg& r, const Slot& s, Acc& c)
    {
        AJ::interact(r.a, s, c.a);
        if constexpr (kSplit) {
            Slot s2 = s;
            s2.v[3] *= 2.0f;
            TS::interact(r.t, s2, c.t);
        } else {
            TS::interact(r.t, s, c.t);
        }
    }
    __device__ static void flush(Acc& c)
    {
        AJ::flush(c.a);
        TS::flush(c.t);
    }
    // split32: each interaction's lane sums evaluated again in FP64 only if
    // they are not finite, as in its own kernel
    __device__ static void fallback_parts(const Args& a, int i, int j0, int j1, int jstep,
                                          double* tot)
    {
        if constexpr (kSplit) {
            double t1[AJ::kNacc], t2[TS::kNacc];
            for (int k = 0; k < AJ::kNacc; ++k) t1[k] = tot[k];
            for (int k = 0; k < TS::kNacc; ++k) t2[k] = tot[AJ::kNacc + k];
            if (!all_finite(t1)) fallback64<typename AJ::Fallback64>(a, i, j0, j1, jstep, tot);
            if (!all_finite(t2))
                fallback64<typename TS::Fallback64>(a, i, j0, j1, jstep, tot + AJ::kNacc);
        }
    }
    __device__ static double get(const Acc& c, int k)
    {
        return k < AJ::kNacc ? AJ::get(c.a, k) : TS::get(c.t, k - AJ::kNacc);
    }
    __device__ static void combine(double* t, const double* p)
    {
        AJ::combine(t, p);
        TS::combine(t + AJ::kNacc, p + AJ::kNacc);
    }
    __device__ static void store(const Args& a, int i, const double* t)
    {
        AJ::store(a, i, t);
        Args b = a;
        b.out[0] = a.out[6];
        b.out[1] = a.out[7];
        TS::store(b, i, t + AJ::kNacc);
    }
};

// The step size (block_tau_kernel's work) and then the order-4 predictor
// of every particle into the r0, v0 scratch from the start forces a0, j0,
// in one workgroup: the predictor that stage 0 runs in its epilogue, for a
// step whose start forces came with Tstep in one pass.
__global__ __launch_bounds__(kTauThreads) void tau_predict4_kernel(int n, HermiteArrays h,
                                                                   int adaptive, double* ctl)
{
    __shared__ double red[kTauThreads];
    __shared__ double tau_sh;
    block_tau_run(n, h.ts, adaptive, ctl, red);
    if (threadIdx.x == 0) tau_sh = ctl[kCtlTau];
    __syncthreads();
    const double tau = tau_sh;
    if (tau == 0) return;
    for (int i = static_cast<int>(threadIdx.x); i < n; i += kTauThreads) {
#pragma unroll
        for (int d = 0; d < 3; ++d)
            predict_dim<4>(h, i, d, tau, h.a0[d][i], h.j0[d][i], nullptr, nullptr, h.r0, h.v0);
    }
}

// Whether an order-4 step of n particles in `mode` takes the fused launches:
// the small-set kernel's range (as pairwise_call decides it for AccJerk).
bool fused4(int mode, int n)
{
    return mode != TUPAN_MODE_IEEE64 && n > kSmallMinJ && n <= kSmallMaxJ &&
           static_cast<long long>(n) * n <= kSmallPairs;
}

template <class Base, int Stage>
int force_stage4(int n, const HermiteArrays& h, double* ctl, int adaptive, hipStream_t s)
{
    using Op = OpHermite4<Base, Stage>;
    typename Op::Args a{};
    double* const* R = Stage == 0 ? h.r : (Stage == 1 ? h.r0 : h.s0);
    double* const* V = Stage == 0 ? h.v : (Stage == 1 ? h.v0 : h.c0);
    const double* in[8] = {h.m, R[0], R[1], R[2], h.e2, V[0], V[1], V[2]};
    double* const* A = Stage == 0 ? h.a0 : h.a1;
    double* const* J = Stage == 0 ? h.j0 : h.j1;
    for (int f = 0; f < 8; ++f) {
        a.ip[f] = in[f];
        a.jp[f] = in[f];
    }
    for (int k = 0; k < 3; ++k) {
        a.out[k] = A[k];
        a.out[3 + k] = J[k];
    }
    a.ni = n;
    a.nj = n;
    a.nsplit = 1;
    a.jchunk = n;
    a.h = h;
    a.ctl = ctl;
    a.adaptive = adaptive;
    constexpr int ipb = kSmallBlock / kSmallGroups;
    const unsigned grid = static_cast<unsigned>((n + ipb - 1) / ipb);
    hipLaunchKernelGGL((pair_small_kernel<Op, kSmallBlock, kSmallGroups, small_tile<Base>()>),
                       dim3(grid), dim3(kSmallBlock), 0, s, a);
    return check_launch("hermite force stage");
}

template <class Base>
int fused_step4(int n, const HermiteArrays& h, double* ctl, int adaptive, hipStream_t s)
{
    int rc = force_stage4<Base, 0>(n, h, ctl, adaptive, s);
    if (rc == TUPAN_OK) rc = force_stage4<Base, 1>(n, h, ctl, adaptive, s);
    if (rc == TUPAN_OK) rc = force_stage4<Base, 2>(n, h, ctl, adaptive, s);
    return rc;
}

// Adaptive order-4 step of a small set in four launches: Tstep and the
// start forces in one pass, the step size and the predictor, then stages 1
// and 2 (the same operations as the five-launch form, the same results).
template <class AJ, class TS>
int fused_adaptive_step4(int n, const HermiteArrays& h, double eta, double* ctl, hipStream_t s)
{
    using Op = OpAccJerkTstep<AJ, TS>;
    typename Op::Args a{};
    const double* in[8] = {h.m, h.r[0], h.r[1], h.r[2], h.e2, h.v[0], h.v[1], h.v[2]};
    for (int f = 0; f < 8; ++f) {
        a.ip[f] = in[f];
        a.jp[f] = in[f];
    }
    for (int k = 0; k < 3; ++k) {
        a.out[k] = h.a0[k];
        a.out[3 + k] = h.j0[k];
    }
    a.out[6] = h.ts;
    a.out[7] = h.tsij;
    a.ni = n;
    a.nj = n;
    a.nsplit = 1;
    a.jchunk = n;
    a.eta = eta;
    constexpr int ipb = kSmallBlock / kSmallGroups;
    const unsigned grid = static_cast<unsigned>((n + ipb - 1) / ipb);
    hipLaunchKernelGGL((pair_small_kernel<Op, kSmallBlock, kSmallGroups, small_tile<Op>()>),
                       dim3(grid), dim3(kSmallBlock), 0, s, a);
    int rc = check_launch("hermite start forces and Tstep");
    if (rc != TUPAN_OK) return rc;
    hipLaunchKernelGGL(tau_predict4_kernel, dim3(1), dim3(kTauThreads), 0, s, n, h, 1, ctl);
    rc = check_launch("hermite tau and predictor");
    if (rc == TUPAN_OK) rc = force_stage4<AJ, 1>(n, h, ctl, 1, s);
    if (rc == TUPAN_OK) rc = force_stage4<AJ, 2>(n, h, ctl, 1, s);
    return rc;
}

int forces(int order, int mode, size_t n, const HermiteArrays& h, bool start, hipStream_t s)
{
    // AccJerk: i and j fields m rx ry rz e2 vx vy vz
    const double* in[8] = {h.m, h.r[0], h.r[1], h.r[2], h.e2, h.v[0], h.v[1], h.v[2]};
    double* const* A = start ? h.a0 : h.a1;
    double* const* J = start ? h.j0 : h.j1;
    double* aj[6] = {A[0], A[1], A[2], J[0], J[1], J[2]};
    const int kernel = order == 2 ? K_ACC : K_ACC_JERK;
    const double* in5[5] = {h.m, h.r[0], h.r[1], h.r[2], h.e2};
    int rc = kernel == K_ACC
                 ? pairwise_call<double>(K_ACC, mode, n, in5, n, in5, aj, 0.0, s)
                 : pairwise_call<double>(K_ACC_JERK, mode, n, in, n, in, aj, 0.0, s);
    if (rc != TUPAN_OK || order < 6) return rc;
    double* const* S = start ? h.s0 : h.s1;
    double* const* C = start ? h.c0 : h.c1;
    const double* in14[14] = {h.m,    h.r[0], h.r[1], h.r[2], h.e2,   h.v[0], h.v[1],
                              h.v[2], A[0],   A[1],   A[2],   J[0],   J[1],   J[2]};
    double* sc[6] = {S[0], S[1], S[2], C[0], C[1], C[2]};
    const int smode = mode == TUPAN_MODE_SPLIT32 ? TUPAN_MODE_FP64 : mode;
    return pairwise_call<double>(K_SNAP_CRACKLE, smode, n, in14, n, in14, sc, 0.0, s);
}

template <int Order>
int step_order(int adaptive, int mode, int n, const HermiteArrays& h, double eta, double* ctl,
               hipStream_t s)
{
    const unsigned grid = static_cast<unsigned>((n + 255) / 256);
    if constexpr (Order == 4) {
        if (adaptive && fused4(mode, n)) {
            return mode == TUPAN_MODE_SPLIT32
                       ? fused_adaptive_step4<OpSplitAccJerk, OpSplitTstep>(n, h, eta, ctl, s)
                       : fused_adaptive_step4<OpAccJerk<Fast64>, OpTstep<Fast64>>(n, h, eta,
                                                                                  ctl, s);
        }
    }
    if (adaptive) {
        const double* in[8] = {h.m, h.r[0], h.r[1], h.r[2], h.e2, h.v[0], h.v[1], h.v[2]};
        double* out[2] = {h.ts, h.tsij};
        int rc = pairwise_call<double>(K_TSTEP, mode, n, in, n, in, out, eta, s);
        if (rc != TUPAN_OK) return rc;
    }
    hipLaunchKernelGGL(block_tau_kernel, dim3(1), dim3(kTauThreads), 0, s, n, h.ts, adaptive,
                       ctl);
    int rc = check_launch("hermite_tau");
    if (rc != TUPAN_OK) return rc;
    if constexpr (Order == 4) {
        if (fused4(mode, n)) {
            return mode == TUPAN_MODE_SPLIT32
                       ? fused_step4<OpSplitAccJerk>(n, h, ctl, adaptive, s)
                       : fused_step4<OpAccJerk<Fast64>>(n, h, ctl, adaptive, s);
        }
    }
    rc = forces(Order, mode, n, h, true, s);
    if (rc != TUPAN_OK) return rc;
    hipLaunchKernelGGL((hermite_predict_kernel<Order>), dim3(grid), dim3(256), 0, s, n, h, ctl);
    rc = check_launch("hermite_predict");
    for (int k = 0; k < 2 && rc == TUPAN_OK; ++k) {
        rc = forces(Order, mode, n, h, false, s);
        if (rc != TUPAN_OK) break;
        hipLaunchKernelGGL((hermite_correct_kernel<Order>), dim3(grid), dim3(256), 0, s, n, h,
                           ctl);
        rc = check_launch("hermite_correct");
    }
    if (rc != TUPAN_OK) return rc;
    hipLaunchKernelGGL(hermite_end_kernel, dim3(grid), dim3(256), 0, s, n, h, adaptive, ctl);
    return check_launch("hermite_end");
}

}  // namespace

int hermite_step_call(int order, int adaptive, int mode, size_t n, void* const* arrays,
                      double* scratch, double* control, double eta, void* stream)
{
    if (order != 2 && order != 4 && order != 6 && order != 8)
        return set_error(TUPAN_ERR_ARG, "Hermite order must be 2, 4, 6 or 8");
    if (mode != TUPAN_MODE_FP64 && mode != TUPAN_MODE_SPLIT32 && mode != TUPAN_MODE_IEEE64)
        return set_error(TUPAN_ERR_ARG, "unknown force mode");
    if (n == 0) return TUPAN_OK;
    if (n > static_cast<size_t>(INT32_MAX)) return set_error(TUPAN_ERR_ARG, "n exceeds 2^31-1");
    if (!arrays || !scratch || !control) return set_error(TUPAN_ERR_ARG, "null argument");
    const int nreq = order >= 6 ? TUPAN_H_NROLES : TUPAN_H_SX;
    for (int q = 0; q < TUPAN_H_NROLES; ++q)
        if ((q < nreq || q >= TUPAN_H_TSTEP) && !arrays[q])
            return set_error(TUPAN_ERR_ARG, "null particle array");
    HermiteArrays h{};
    auto d = [&](int q) { return static_cast<double*>(arrays[q]); };
    h.m = d(TUPAN_H_MASS);
    h.e2 = d(TUPAN_H_EPS2);
    for (int k = 0; k < 3; ++k) {
        h.r[k] = d(TUPAN_H_RX + k);
        h.v[k] = d(TUPAN_H_VX + k);
        h.a[k] = d(TUPAN_H_AX + k);
        h.j[k] = d(TUPAN_H_JX + k);
        h.s[k] = order >= 6 ? d(TUPAN_H_SX + k) : nullptr;
        h.c[k] = order >= 6 ? d(TUPAN_H_CX + k) : nullptr;
    }
    h.tstep = d(TUPAN_H_TSTEP);
    h.tstepij = d(TUPAN_H_TSTEPIJ);
    h.time = d(TUPAN_H_TIME);
    h.nstep = static_cast<long long*>(arrays[TUPAN_H_NSTEP]);
    double* sp = scratch;
    for (int k = 0; k < 3; ++k) {
        h.r0[k] = sp + (0 + k) * n;
        h.v0[k] = sp + (3 + k) * n;
        h.a0[k] = sp + (6 + k) * n;
        h.j0[k] = sp + (9 + k) * n;
        h.s0[k] = sp + (12 + k) * n;
        h.c0[k] = sp + (15 + k) * n;
        h.a1[k] = sp + (20 + k) * n;
        h.j1[k] = sp + (23 + k) * n;
        h.s1[k] = sp + (26 + k) * n;
        h.c1[k] = sp + (29 + k) * n;
    }
    h.ts = sp + 18 * n;
    h.tsij = sp + 19 * n;
    hipStream_t s = static_cast<hipStream_t>(stream);
    const int ni = static_cast<int>(n);
    switch (order) {
    case 2: return step_order<2>(adaptive, mode, ni, h, eta, control, s);
    case 4: return step_order<4>(adaptive, mode, ni, h, eta, control, s);
    case 6: return step_order<6>(adaptive, mode, ni, h, eta, control, s);
    default: return step_order<8>(adaptive, mode, ni, h, eta, control, s);
    }
}

}  // namespace tupan
