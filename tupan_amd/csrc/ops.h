// Pair interactions of the all-pairs kernels (Phi, Acc, AccJerk,
// SnapCrackle, Tstep, the NREG sums and the split-FP32 forms), as policies
// for the tile engine (engine.h); the post-Newtonian pair acceleration is in
// pn.h.  Shared by pairwise.hip and sia.hip, which must
// evaluate them with the same operation order.
//
// Per-pair mathematics follows the reference kernel cores
//   phi_kernel_common.h:7-32, acc_kernel_common.h:7-38,
//   acc_jerk_kernel_common.h:7-56, snap_crackle_kernel_common.h:7-95,
//   tstep_kernel_common.h:7-60, with the softening of smoothing.h:10-186;
// the loop structure is the LDS tile engine of engine.h.
//
// Arithmetic modes (Inv policies):
//   Fast64  FP64; 1/R from v_rsq_f64 plus a second-order Newton step, FMA
//           contraction.  The production FP64 path.
//   Ieee64  FP64; 1/R^2 by IEEE division and 1/R by IEEE square root in the
//           reference's operation order, no contraction.  With one j split
//           its per-i sums equal the reference C backend bit for bit.
//   Fast32  FP32 throughout (the reference's float build, `--use_sp`).
// plus the split-FP32 interactions (SplitAcc, SplitAccJerk, SplitTstep): FP64
// inputs and outputs, pair arithmetic in FP32 on hi+lo coordinates relative to
// the workgroup's first i-particle, 1/R from the v_rsq_f32 estimate, FP32 partial
// sums flushed into FP64 every kFlush j-particles.
//
// This file is compiled with -ffp-contract=off: every fused multiply-add
// below is explicit.
#pragma once

#include "engine.h"

namespace tupan {

// ---------------------------------------------------------------------------
// Inverse-power policies.
// ---------------------------------------------------------------------------
struct Fast64 {
    using T = double;
    static constexpr bool kFused = true;
    // q1 = 1/R, q2 = 1/R^2 with R^2 = r2 + e2; 0 for masked pairs.
    __device__ static __forceinline__ void q12(T r2, T e2, bool mask, T& q1, T& q2)
    {
        T y = rsqrt_refined(r2 + e2);
        y = mask ? y : T(0);
        q1 = y;
        q2 = y * y;
    }
    __device__ static __forceinline__ T eta_over_sqrt(T eta, T w2)
    {
        return eta * rsqrt_refined(w2);
    }
};

struct Fast32 {
    using T = float;
    static constexpr bool kFused = true;
    __device__ static __forceinline__ void q12(T r2, T e2, bool mask, T& q1, T& q2)
    {
        T y = rsqrt_refined(r2 + e2);
        y = mask ? y : T(0);
        q1 = y;
        q2 = y * y;
    }
    __device__ static __forceinline__ T eta_over_sqrt(T eta, T w2)
    {
        return eta * rsqrt_refined(w2);
    }
};

struct Ieee64 {
    using T = double;
    static constexpr bool kFused = false;
    __device__ static __forceinline__ void q12(T r2, T e2, bool mask, T& q1, T& q2)
    {
        T q = 1.0 / (r2 + e2);
        q = mask ? q : 0.0;
        q2 = q;
        q1 = __builtin_sqrt(q);
    }
    __device__ static __forceinline__ T eta_over_sqrt(T eta, T w2)
    {
        return eta / __builtin_sqrt(w2);
    }
};

// a*b + c, fused or not according to the policy.
template <class P, class T>
__device__ __forceinline__ T mad(T a, T b, T c)
{
    if constexpr (P::kFused) {
        return fma_t(a, b, c);
    } else {
        return a * b + c;
    }
}

// c - a*b
template <class P, class T>
__device__ __forceinline__ T nmad(T a, T b, T c)
{
    if constexpr (P::kFused) {
        return fma_t(-a, b, c);
    } else {
        return c - a * b;
    }
}

// x*x + y*y + z*z in the reference's order when not fused.
template <class P, class T>
__device__ __forceinline__ T dot3(T ax, T ay, T az, T bx, T by, T bz)
{
    if constexpr (P::kFused) {
        return fma_t(az, bz, fma_t(ay, by, ax * bx));
    } else {
        return ax * bx + ay * by + az * bz;
    }
}

struct NoFrame {};

template <class T, int N>
struct alignas(16) Slots {
    T v[N];
};

// Argument block shared by all pairwise ops.
template <class T>
struct PairArgs {
    const T* ip[14];
    const T* jp[14];
    T* out[8];
    int ni, nj;
    int nsplit, jchunk;
    double* partial;
    T eta;
    T clight[5];  // PNAcc: 1/c^2, 1/c^4, 1/c^5, 1/c^6, 1/c^7
    int order;    // PNAcc: post-Newtonian order
    int exact;    // host-side launch planning: keep whole-tile j-chunks
    // split AccJerk: 0 here when no i and j share a position with different
    // velocities (pairwise.hip: the coincidence scan); null: not scanned
    const int* coincide;
};

// ---------------------------------------------------------------------------
// Phi.  i: rx ry rz e2 ; j: m rx ry rz e2 (padded to 6).
// ---------------------------------------------------------------------------
template <class P>
struct OpPhi {
    using T = typename P::T;
    using Args = PairArgs<T>;
    using Frame = NoFrame;
    using Slot = Slots<T, (sizeof(T) == 8 ? 6 : 8)>;
    struct IReg { T x, y, z, e2; };
    struct Acc { T phi; };
    static constexpr int kNacc = 1;
    static constexpr int kFlush = 1 << 30;
    // uniform softening: when every i of the workgroup and every j of a tile
    // carry the same e2, the engine passes the pair value e2 + e2 once
    static constexpr bool kUniformE2 = true;
    __device__ static T e2_of_i(const Args& a, int i) { return a.ip[4][i]; }
    __device__ static T slot_e2(const Slot& s) { return s.v[4]; }

    __device__ static Frame frame(const Args&, int) { return {}; }
    __device__ static void load_i(const Args& a, const Frame&, int i, IReg& r)
    {
        r.x = a.ip[1][i]; r.y = a.ip[2][i]; r.z = a.ip[3][i]; r.e2 = a.ip[4][i];
    }
    __device__ static void load_j(const Args& a, const Frame&, int j, Slot& s)
    {
#pragma unroll
        for (int f = 0; f < 5; ++f) s.v[f] = a.jp[f][j];
    }
    __device__ static void zero(Acc& c) { c.phi = 0; }
    template <bool kU = false>
    __device__ static __forceinline__ void interact(const IReg& r, const Slot& s, Acc& c,
                                                    T e2u = T(0))
    {
        T dx = r.x - s.v[1];
        T dy = r.y - s.v[2];
        T dz = r.z - s.v[3];
        T e2 = kU ? e2u : r.e2 + s.v[4];
        T r2 = dot3<P>(dx, dy, dz, dx, dy, dz);
        T q1, q2;
        P::q12(r2, e2, r2 > T(0), q1, q2);
        c.phi = nmad<P>(s.v[0], q1, c.phi);
    }
    __device__ static void flush(Acc&) {}
    __device__ static double get(const Acc& c, int) { return static_cast<double>(c.phi); }
    __device__ static void combine(double* t, const double* p) { t[0] += p[0]; }
    __device__ static void store(const Args& a, int i, const double* t)
    {
        a.out[0][i] = static_cast<T>(t[0]);
    }
};

// ---------------------------------------------------------------------------
// Acc.  i: rx ry rz e2 ; j: m rx ry rz e2.
// ---------------------------------------------------------------------------
template <class P>
struct OpAcc {
    using T = typename P::T;
    using Args = PairArgs<T>;
    using Frame = NoFrame;
    using Slot = Slots<T, (sizeof(T) == 8 ? 6 : 8)>;
    struct IReg { T x, y, z, e2; };
    struct Acc { T ax, ay, az; };
    static constexpr int kNacc = 3;
    static constexpr int kFlush = 1 << 30;
    // uniform softening: when every i of the workgroup and every j of a tile
    // carry the same e2, the engine passes the pair value e2 + e2 once
    static constexpr bool kUniformE2 = true;
    __device__ static T e2_of_i(const Args& a, int i) { return a.ip[4][i]; }
    __device__ static T slot_e2(const Slot& s) { return s.v[4]; }

    __device__ static Frame frame(const Args&, int) { return {}; }
    __device__ static void load_i(const Args& a, const Frame&, int i, IReg& r)
    {
        r.x = a.ip[1][i]; r.y = a.ip[2][i]; r.z = a.ip[3][i]; r.e2 = a.ip[4][i];
    }
    __device__ static void load_j(const Args& a, const Frame&, int j, Slot& s)
    {
#pragma unroll
        for (int f = 0; f < 5; ++f) s.v[f] = a.jp[f][j];
    }
    // Unmasked uniform tiles (the FP64 fast path): a pair at zero
    // separation has r2 = 0 and R^2 = e2u exactly, and adds -q3 * 0 = 0
    // when its q3 = m_j / e2u^1.5 is finite (self_safe, per tile), so the
    // tile runs without the r2 > 0 test and gives the masked sums.
    static constexpr bool kUnmask = std::is_same<P, Fast64>::value;
    __device__ static bool unmask_allowed(const Args&) { return true; }
    __device__ static T self_y(T e2u)
    {
        T q1, q2;
        P::q12(T(0), e2u, true, q1, q2);
        return q1;
    }
    __device__ static bool self_safe(T y0, const Slot& s)
    {
        const T q3 = ((y0 * y0) * y0) * s.v[0];
        return __builtin_isfinite(q3);
    }
    __device__ static void zero(Acc& c) { c.ax = c.ay = c.az = 0; }
    template <bool kU = false, bool kM = true>
    __device__ static __forceinline__ void interact(const IReg& r, const Slot& s, Acc& c,
                                                    T e2u = T(0))
    {
        T dx = r.x - s.v[1];
        T dy = r.y - s.v[2];
        T dz = r.z - s.v[3];
        T e2 = kU ? e2u : r.e2 + s.v[4];
        T r2 = dot3<P>(dx, dy, dz, dx, dy, dz);
        T q1, q2;
        P::q12(r2, e2, (kM || !kU) ? r2 > T(0) : true, q1, q2);
        T q3 = q2 * q1;
        q3 *= s.v[0];
        c.ax = nmad<P>(q3, dx, c.ax);
        c.ay = nmad<P>(q3, dy, c.ay);
        c.az = nmad<P>(q3, dz, c.az);
    }
    __device__ static void flush(Acc&) {}
    __device__ static double get(const Acc& c, int k)
    {
        return static_cast<double>(k == 0 ? c.ax : (k == 1 ? c.ay : c.az));
    }
    __device__ static void combine(double* t, const double* p)
    {
        t[0] += p[0]; t[1] += p[1]; t[2] += p[2];
    }
    __device__ static void store(const Args& a, int i, const double* t)
    {
#pragma unroll
        for (int k = 0; k < 3; ++k) a.out[k][i] = static_cast<T>(t[k]);
    }
};

// ---------------------------------------------------------------------------
// AccJerk.  i: rx ry rz e2 vx vy vz ; j: m rx ry rz e2 vx vy vz.
// ---------------------------------------------------------------------------
template <class P>
struct OpAccJerk {
    using T = typename P::T;
    using Args = PairArgs<T>;
    using Frame = NoFrame;
    using Slot = Slots<T, 8>;
    struct IReg { T x, y, z, e2, vx, vy, vz; };
    struct Acc { T ax, ay, az, jx, jy, jz; };
    static constexpr int kNacc = 6;
    static constexpr int kFlush = 1 << 30;
    // uniform softening: when every i of the workgroup and every j of a tile
    // carry the same e2, the engine passes the pair value e2 + e2 once
    static constexpr bool kUniformE2 = true;
    __device__ static T e2_of_i(const Args& a, int i) { return a.ip[4][i]; }
    __device__ static T slot_e2(const Slot& s) { return s.v[4]; }

    __device__ static Frame frame(const Args&, int) { return {}; }
    __device__ static void load_i(const Args& a, const Frame&, int i, IReg& r)
    {
        r.x = a.ip[1][i]; r.y = a.ip[2][i]; r.z = a.ip[3][i]; r.e2 = a.ip[4][i];
        r.vx = a.ip[5][i]; r.vy = a.ip[6][i]; r.vz = a.ip[7][i];
    }
    __device__ static void load_j(const Args& a, const Frame&, int j, Slot& s)
    {
#pragma unroll
        for (int f = 0; f < 8; ++f) s.v[f] = a.jp[f][j];
    }
    // Unmasked uniform tiles (the FP64 fast path), as OpSplitAccJerkT's: the
    // call's coincidence scan (pairwise.hip) clears the flag only when no two
    // particles share a position with different velocities and every j's
    // zero-separation term is finite.
    static constexpr bool kCoincideScan = std::is_same<P, Fast64>::value;
    static constexpr bool kUnmask = kCoincideScan;
    __device__ static bool unmask_allowed(const Args& a)
    {
        return a.coincide != nullptr && *a.coincide == 0;
    }
    __device__ static bool self_safe_j(const Args& a, int j)
    {
        const T e2 = a.jp[4][j];
        T q1, q2;
        P::q12(T(0), e2 + e2, true, q1, q2);
        const T q3 = (q2 * q1) * a.jp[0][j];
        return __builtin_isfinite(q2) && __builtin_isfinite(q3);
    }
    __device__ static void zero(Acc& c) { c.ax = c.ay = c.az = c.jx = c.jy = c.jz = 0; }
    template <bool kU = false, bool kM = true>
    __device__ static __forceinline__ void interact(const IReg& r, const Slot& s, Acc& c,
                                                    T e2u = T(0))
    {
        T dx = r.x - s.v[1];
        T dy = r.y - s.v[2];
        T dz = r.z - s.v[3];
        T e2 = kU ? e2u : r.e2 + s.v[4];
        T ux = r.vx - s.v[5];
        T uy = r.vy - s.v[6];
        T uz = r.vz - s.v[7];
        T r2 = dot3<P>(dx, dy, dz, dx, dy, dz);
        T rv = dot3<P>(dx, dy, dz, ux, uy, uz);
        T q1, q2;
        P::q12(r2, e2, (kM || !kU) ? r2 > T(0) : true, q1, q2);
        T q3 = q2 * q1;
        T alpha = T(3) * rv * q2;
        ux = nmad<P>(alpha, dx, ux);
        uy = nmad<P>(alpha, dy, uy);
        uz = nmad<P>(alpha, dz, uz);
        q3 *= s.v[0];
        c.ax = nmad<P>(q3, dx, c.ax);
        c.ay = nmad<P>(q3, dy, c.ay);
        c.az = nmad<P>(q3, dz, c.az);
        c.jx = nmad<P>(q3, ux, c.jx);
        c.jy = nmad<P>(q3, uy, c.jy);
        c.jz = nmad<P>(q3, uz, c.jz);
    }
    __device__ static void flush(Acc&) {}
    __device__ static double get(const Acc& c, int k)
    {
        const T v[6] = {c.ax, c.ay, c.az, c.jx, c.jy, c.jz};
        return static_cast<double>(v[k]);
    }
    __device__ static void combine(double* t, const double* p)
    {
#pragma unroll
        for (int k = 0; k < 6; ++k) t[k] += p[k];
    }
    __device__ static void store(const Args& a, int i, const double* t)
    {
#pragma unroll
        for (int k = 0; k < 6; ++k) a.out[k][i] = static_cast<T>(t[k]);
    }
};

// ---------------------------------------------------------------------------
// SnapCrackle.  i: rx ry rz e2 vx vy vz ax ay az jx jy jz ;
//               j: m rx ry rz e2 vx vy vz ax ay az jx jy jz (padded to 16).
// ---------------------------------------------------------------------------
template <class P>
struct OpSnapCrackle {
    using T = typename P::T;
    using Args = PairArgs<T>;
    using Frame = NoFrame;
    using Slot = Slots<T, 16>;
    struct IReg { T v[13]; };
    struct Acc { T s[6]; };
    static constexpr int kNacc = 6;
    static constexpr int kFlush = 1 << 30;

    __device__ static Frame frame(const Args&, int) { return {}; }
    __device__ static void load_i(const Args& a, const Frame&, int i, IReg& r)
    {
#pragma unroll
        for (int f = 0; f < 13; ++f) r.v[f] = a.ip[f + 1][i];
    }
    __device__ static void load_j(const Args& a, const Frame&, int j, Slot& s)
    {
#pragma unroll
        for (int f = 0; f < 14; ++f) s.v[f] = a.jp[f][j];
    }
    __device__ static void zero(Acc& c)
    {
#pragma unroll
        for (int k = 0; k < 6; ++k) c.s[k] = 0;
    }
    __device__ static __forceinline__ void interact(const IReg& r, const Slot& s, Acc& c)
    {
        T dx = r.v[0] - s.v[1];
        T dy = r.v[1] - s.v[2];
        T dz = r.v[2] - s.v[3];
        T e2 = r.v[3] + s.v[4];
        T ux = r.v[4] - s.v[5];
        T uy = r.v[5] - s.v[6];
        T uz = r.v[6] - s.v[7];
        T Ax = r.v[7] - s.v[8];
        T Ay = r.v[8] - s.v[9];
        T Az = r.v[9] - s.v[10];
        T Jx = r.v[10] - s.v[11];
        T Jy = r.v[11] - s.v[12];
        T Jz = r.v[12] - s.v[13];
        T r2 = dot3<P>(dx, dy, dz, dx, dy, dz);
        T rv = dot3<P>(dx, dy, dz, ux, uy, uz);
        T v2 = dot3<P>(ux, uy, uz, ux, uy, uz);
        T rJ = dot3<P>(dx, dy, dz, Jx, Jy, Jz);
        T rA = dot3<P>(dx, dy, dz, Ax, Ay, Az);
        T vA = dot3<P>(ux, uy, uz, Ax, Ay, Az);
        T q1, q2;
        P::q12(r2, e2, r2 > T(0), q1, q2);
        T q3 = q2 * q1;

        T al = rv * q2;
        T al2 = al * al;
        T be = T(3) * mad<P>(v2 + rA, q2, al2);
        T ga = mad<P>(mad<P>(T(3), vA, rJ), q2, al * nmad<P>(T(4), al2, be));
        al *= T(3);
        ga *= T(3);
        ux = nmad<P>(al, dx, ux);
        uy = nmad<P>(al, dy, uy);
        uz = nmad<P>(al, dz, uz);
        T tw = T(2) * al;
        Ax -= mad<P>(tw, ux, be * dx);
        Ay -= mad<P>(tw, uy, be * dy);
        Az -= mad<P>(tw, uz, be * dz);
        al *= T(3);
        be *= T(3);
        Jx -= mad<P>(ga, dx, mad<P>(al, Ax, be * ux));
        Jy -= mad<P>(ga, dy, mad<P>(al, Ay, be * uy));
        Jz -= mad<P>(ga, dz, mad<P>(al, Az, be * uz));
        q3 *= s.v[0];
        c.s[0] = nmad<P>(q3, Ax, c.s[0]);
        c.s[1] = nmad<P>(q3, Ay, c.s[1]);
        c.s[2] = nmad<P>(q3, Az, c.s[2]);
        c.s[3] = nmad<P>(q3, Jx, c.s[3]);
        c.s[4] = nmad<P>(q3, Jy, c.s[4]);
        c.s[5] = nmad<P>(q3, Jz, c.s[5]);
    }
    __device__ static void flush(Acc&) {}
    __device__ static double get(const Acc& c, int k) { return static_cast<double>(c.s[k]); }
    __device__ static void combine(double* t, const double* p)
    {
#pragma unroll
        for (int k = 0; k < 6; ++k) t[k] += p[k];
    }
    __device__ static void store(const Args& a, int i, const double* t)
    {
#pragma unroll
        for (int k = 0; k < 6; ++k) a.out[k][i] = static_cast<T>(t[k]);
    }
};

// ---------------------------------------------------------------------------
// Tstep.  i: m rx ry rz e2 vx vy vz ; j: the same.  Accumulators: the sum
// and the max (starting at 0) of the pair frequencies w^2; outputs
// dt_a = eta/sqrt(1 + sum), dt_b = eta/sqrt(1 + max) (tstep_kernel.c:45-46).
// ---------------------------------------------------------------------------
template <class P>
struct OpTstep {
    using T = typename P::T;
    using Args = PairArgs<T>;
    using Frame = NoFrame;
    using Slot = Slots<T, 8>;
    struct IReg { T m, x, y, z, e2, vx, vy, vz, eta; };
    struct Acc { T sum, mx; };
    static constexpr int kNacc = 2;
    static constexpr int kFlush = 1 << 30;
    // uniform softening: when every i of the workgroup and every j of a tile
    // carry the same e2, the engine passes the pair value e2 + e2 once
    static constexpr bool kUniformE2 = true;
    __device__ static T e2_of_i(const Args& a, int i) { return a.ip[4][i]; }
    __device__ static T slot_e2(const Slot& s) { return s.v[4]; }

    __device__ static Frame frame(const Args&, int) { return {}; }
    __device__ static void load_i(const Args& a, const Frame&, int i, IReg& r)
    {
        r.m = a.ip[0][i]; r.x = a.ip[1][i]; r.y = a.ip[2][i]; r.z = a.ip[3][i];
        r.e2 = a.ip[4][i]; r.vx = a.ip[5][i]; r.vy = a.ip[6][i]; r.vz = a.ip[7][i];
        r.eta = a.eta;
    }
    __device__ static void load_j(const Args& a, const Frame&, int j, Slot& s)
    {
#pragma unroll
        for (int f = 0; f < 8; ++f) s.v[f] = a.jp[f][j];
    }
    __device__ static void zero(Acc& c) { c.sum = 0; c.mx = 0; }
    template <bool kU = false>
    __device__ static __forceinline__ void interact(const IReg& r, const Slot& s, Acc& c,
                                                    T e2u = T(0))
    {
        T dx = r.x - s.v[1];
        T dy = r.y - s.v[2];
        T dz = r.z - s.v[3];
        T e2 = kU ? e2u : r.e2 + s.v[4];
        T ux = r.vx - s.v[5];
        T uy = r.vy - s.v[6];
        T uz = r.vz - s.v[7];
        T mt = r.m + s.v[0];
        T r2 = dot3<P>(dx, dy, dz, dx, dy, dz);
        T rv = dot3<P>(dx, dy, dz, ux, uy, uz);
        T v2 = dot3<P>(ux, uy, uz, ux, uy, uz);
        const bool mask = r2 > T(0);
        T q1, q2;
        // a masked pair's term is zeroed at the end, so the fast policies
        // skip masking 1/R here (the reference-order policy keeps it)
        P::q12(r2, e2, P::kFused ? true : mask, q1, q2);
        // a = 2, b = (1 + a/2) * q2 = 2 q2
        T b = T(2) * q2;
        T phi = mt * q1;
        T w2 = q2 * mad<P>(T(2), phi, v2);
        T ga = q2 * mad<P>(b, phi, w2);
        ga *= P::eta_over_sqrt(r.eta, w2);
        w2 = nmad<P>(ga, rv, w2);
        w2 = mask ? w2 : T(0);
        c.sum += w2;
        c.mx = fmax_t(w2, c.mx);
    }
    __device__ static void flush(Acc&) {}
    __device__ static double get(const Acc& c, int k)
    {
        return static_cast<double>(k == 0 ? c.sum : c.mx);
    }
    __device__ static void combine(double* t, const double* p)
    {
        t[0] += p[0];
        t[1] = __builtin_fmax(t[1], p[1]);
    }
    __device__ static void store(const Args& a, int i, const double* t)
    {
        const double eta = static_cast<double>(a.eta);
        a.out[0][i] = static_cast<T>(eta / __builtin_sqrt(1.0 + t[0]));
        a.out[1][i] = static_cast<T>(eta / __builtin_sqrt(1.0 + t[1]));
    }
};

// ---------------------------------------------------------------------------
// Split-FP32 Acc and AccJerk.
//
// Frame: the FP64 position and velocity of the workgroup's first i-particle.
// Every coordinate is taken relative to it in FP64 and split into hi + lo
// floats; a pair separation is (hi_i - hi_j) + (lo_i - lo_j), accurate to
// about one FP32 rounding of the separation itself, and exactly 0 for a
// particle paired with itself.  Velocity differences are plain FP32 (relative
// to the frame).  Accumulation: FP32 over kFlush j-particles, then into FP64.
// ---------------------------------------------------------------------------
struct SplitFrame {
    double cx, cy, cz, cvx, cvy, cvz;
};

// Softened 1/R for the split interactions.  R^2 is formed as
// fma(dz, dz, fma(dy, dy, fma(dx, dx, e2))), so e2 enters before any
// rounding: rounding r^2 first and adding e2 afterwards rounds twice, and for
// pairs whose e2 is below half an FP32 ulp of r^2 the second rounding drops e2
// every time, which makes every far force too strong by ~1.5 e2 / R^2 on
// average (+1.2e-8 of the virial on a Plummer sphere with eps = 1e-4) and
// shows up as an energy error proportional to the change of the potential
// energy.  The mask is R^2 > e2: a pair at zero separation (a particle with
// itself, or two coincident particles) gives R^2 == e2 exactly and is
// dropped, as the reference's r2 > 0 drops it; so is a pair whose r^2 is
// below half an ulp of e2 (a separation under 2^-12 of the softening length),
// which the reference keeps.
//
// The raw v_rsq_f32 estimate is within ~1.6 FP32 ulp but not unbiased: over
// every float in [1, 4) (both exponent parities, weight 1/x) its mean relative
// error on gfx950 is kRsqBias.  Accelerations scale as y^3, so the FP32
// partial sums are multiplied by 1 / (1 + 3 kRsqBias) when they are flushed
// into the FP64 sums; with the Newton-refined estimate (unbiased) the factor
// is 1.
constexpr double kRsqBias = -7.35832e-9;

// kMask false: no R^2 > e2 test (the unmasked tiles of the split Acc and
// AccJerk, see their self_safe).
template <bool kRefine, bool kMask = true>
__device__ __forceinline__ float split_rsqrt(float dx, float dy, float dz, float e2)
{
    const float R2 = __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, __builtin_fmaf(dx, dx, e2)));
    const float y = kRefine ? rsqrt_refined(R2) : __builtin_amdgcn_rsqf(R2);
    if constexpr (!kMask) return y;
    return R2 > e2 ? y : 0.0f;
}

// 1/R of a pair at zero separation under the uniform softening e2u: what
// split_rsqrt gives for dx = dy = dz = 0 (its R^2 is then e2u exactly).
template <bool kRefine>
__device__ __forceinline__ float split_self_rsqrt(float e2u)
{
    return kRefine ? rsqrt_refined(e2u) : __builtin_amdgcn_rsqf(e2u);
}

template <bool kRefine>
constexpr double split_scale()
{
    return kRefine ? 1.0 : 1.0 / (1.0 + 3.0 * kRsqBias);
}

template <bool kRefine>
struct OpSplitAccT {
    using Args = PairArgs<double>;
    using Frame = SplitFrame;
    // sums that overflow FP32 are recomputed in FP64 (engine.h: fallback64)
    using Fallback64 = OpAcc<Fast64>;
    // xh yh zh m | xl yl zl e2
    using Slot = Slots<float, 8>;
    struct IReg { float xh, yh, zh, xl, yl, zl, e2; };
    struct Acc { float ax, ay, az; double tx, ty, tz; };
    static constexpr int kNacc = 3;
    static constexpr int kFlush = 64;
    // uniform softening: when every i of the workgroup and every j of a tile
    // carry the same e2, the engine passes the pair value e2 + e2 once
    static constexpr bool kUniformE2 = true;
    __device__ static float e2_of_i(const Args& a, int i) { return static_cast<float>(a.ip[4][i]); }
    __device__ static float slot_e2(const Slot& s) { return s.v[7]; }

    __device__ static Frame frame(const Args& a, int ibase)
    {
        Frame f;
        f.cx = a.ip[1][ibase]; f.cy = a.ip[2][ibase]; f.cz = a.ip[3][ibase];
        f.cvx = f.cvy = f.cvz = 0.0;
        return f;
    }
    __device__ static void load_i(const Args& a, const Frame& f, int i, IReg& r)
    {
        split2(a.ip[1][i] - f.cx, r.xh, r.xl);
        split2(a.ip[2][i] - f.cy, r.yh, r.yl);
        split2(a.ip[3][i] - f.cz, r.zh, r.zl);
        r.e2 = static_cast<float>(a.ip[4][i]);
    }
    __device__ static void load_j(const Args& a, const Frame& f, int j, Slot& s)
    {
        split2(a.jp[1][j] - f.cx, s.v[0], s.v[4]);
        split2(a.jp[2][j] - f.cy, s.v[1], s.v[5]);
        split2(a.jp[3][j] - f.cz, s.v[2], s.v[6]);
        s.v[3] = static_cast<float>(a.jp[0][j]);
        s.v[7] = static_cast<float>(a.jp[4][j]);
    }
    // Unmasked uniform tiles.  The mask R^2 > e2 drops the pairs at zero
    // separation (the reference's r^2 > 0).  Under a uniform softening whose
    // zero-separation term y^3 m_j is finite for every j of the tile
    // (self_safe, checked per tile), such a pair adds -q3 * 0 = 0 to each
    // sum, so the tile runs without the test: the same sums, two
    // instructions fewer per pair.  (A pair with 0 < r^2 below half an FP32
    // ulp of e2, which the mask also drops, then contributes, as in the
    // reference.)
    static constexpr bool kUnmask = true;
    __device__ static bool unmask_allowed(const Args&) { return true; }
    __device__ static float self_y(float e2u) { return split_self_rsqrt<kRefine>(e2u); }
    __device__ static bool self_safe(float y0, const Slot& s)
    {
        const float q3 = (y0 * y0) * (y0 * s.v[3]);
        return __builtin_isfinite(q3);
    }
    __device__ static void zero(Acc& c)
    {
        c.ax = c.ay = c.az = 0.0f;
        c.tx = c.ty = c.tz = 0.0;
    }
    template <bool kU = false, bool kM = true>
    __device__ static __forceinline__ void interact(const IReg& r, const Slot& s, Acc& c,
                                                    float e2u = float(0))
    {
        float dx = (r.xh - s.v[0]) + (r.xl - s.v[4]);
        float dy = (r.yh - s.v[1]) + (r.yl - s.v[5]);
        float dz = (r.zh - s.v[2]) + (r.zl - s.v[6]);
        const float y =
            split_rsqrt<kRefine, kM || !kU>(dx, dy, dz, kU ? e2u : r.e2 + s.v[7]);
        float q3 = (y * y) * (y * s.v[3]);
        c.ax = __builtin_fmaf(-q3, dx, c.ax);
        c.ay = __builtin_fmaf(-q3, dy, c.ay);
        c.az = __builtin_fmaf(-q3, dz, c.az);
    }
    __device__ static void flush(Acc& c)
    {
        constexpr double sc = split_scale<kRefine>();
        c.tx = __builtin_fma(static_cast<double>(c.ax), sc, c.tx);
        c.ty = __builtin_fma(static_cast<double>(c.ay), sc, c.ty);
        c.tz = __builtin_fma(static_cast<double>(c.az), sc, c.tz);
        c.ax = c.ay = c.az = 0.0f;
    }
    __device__ static double get(const Acc& c, int k) { return k == 0 ? c.tx : (k == 1 ? c.ty : c.tz); }
    __device__ static void combine(double* t, const double* p)
    {
        t[0] += p[0]; t[1] += p[1]; t[2] += p[2];
    }
    __device__ static void store(const Args& a, int i, const double* t)
    {
#pragma unroll
        for (int k = 0; k < 3; ++k) a.out[k][i] = t[k];
    }
};

template <bool kRefine>
struct OpSplitAccJerkT {
    using Args = PairArgs<double>;
    using Frame = SplitFrame;
    using Fallback64 = OpAccJerk<Fast64>;
    // xh yh zh m | xl yl zl e2 | vx vy vz pad
    using Slot = Slots<float, 12>;
    struct IReg { float xh, yh, zh, xl, yl, zl, e2, vx, vy, vz; };
    struct Acc { float s[6]; double t[6]; };
    static constexpr int kNacc = 6;
    static constexpr int kFlush = 64;
    // uniform softening: when every i of the workgroup and every j of a tile
    // carry the same e2, the engine passes the pair value e2 + e2 once
    static constexpr bool kUniformE2 = true;
    __device__ static float e2_of_i(const Args& a, int i) { return static_cast<float>(a.ip[4][i]); }
    __device__ static float slot_e2(const Slot& s) { return s.v[7]; }

    __device__ static Frame frame(const Args& a, int ibase)
    {
        Frame f;
        f.cx = a.ip[1][ibase]; f.cy = a.ip[2][ibase]; f.cz = a.ip[3][ibase];
        f.cvx = a.ip[5][ibase]; f.cvy = a.ip[6][ibase]; f.cvz = a.ip[7][ibase];
        return f;
    }
    __device__ static void load_i(const Args& a, const Frame& f, int i, IReg& r)
    {
        split2(a.ip[1][i] - f.cx, r.xh, r.xl);
        split2(a.ip[2][i] - f.cy, r.yh, r.yl);
        split2(a.ip[3][i] - f.cz, r.zh, r.zl);
        r.e2 = static_cast<float>(a.ip[4][i]);
        r.vx = static_cast<float>(a.ip[5][i] - f.cvx);
        r.vy = static_cast<float>(a.ip[6][i] - f.cvy);
        r.vz = static_cast<float>(a.ip[7][i] - f.cvz);
    }
    __device__ static void load_j(const Args& a, const Frame& f, int j, Slot& s)
    {
        split2(a.jp[1][j] - f.cx, s.v[0], s.v[4]);
        split2(a.jp[2][j] - f.cy, s.v[1], s.v[5]);
        split2(a.jp[3][j] - f.cz, s.v[2], s.v[6]);
        s.v[3] = static_cast<float>(a.jp[0][j]);
        s.v[7] = static_cast<float>(a.jp[4][j]);
        s.v[8] = static_cast<float>(a.jp[5][j] - f.cvx);
        s.v[9] = static_cast<float>(a.jp[6][j] - f.cvy);
        s.v[10] = static_cast<float>(a.jp[7][j] - f.cvz);
        s.v[11] = 0.0f;
    }
    // Unmasked uniform tiles, as OpSplitAccT's: a pair at zero separation
    // then has rv = 0, alpha = 0 and adds -q3 * 0 to the accelerations and
    // -q3 * (v_i - v_j) to the jerks, zero for a particle with itself.  Two
    // particles at one position with different velocities would add a
    // non-zero jerk, which the reference drops (r^2 = 0).  The call's
    // coincidence scan (pairwise.hip) clears the flag only when no such
    // pair exists and every j passes self_safe_j, and then every uniform
    // tile goes unmasked.
    static constexpr bool kCoincideScan = true;
    static constexpr bool kUnmask = true;
    __device__ static bool unmask_allowed(const Args& a)
    {
        return a.coincide != nullptr && *a.coincide == 0;
    }
    // j's zero-separation term is finite in any uniform tile it belongs to:
    // there every pair softening is e2_j + e2_j (in FP32, as load_j and
    // e2_of_i round it).
    __device__ static bool self_safe_j(const Args& a, int j)
    {
        const float e2 = static_cast<float>(a.jp[4][j]);
        const float y0 = split_self_rsqrt<kRefine>(e2 + e2);
        const float q2 = y0 * y0;
        const float q3 = q2 * (y0 * static_cast<float>(a.jp[0][j]));
        return __builtin_isfinite(q2) && __builtin_isfinite(q3);
    }
    __device__ static void zero(Acc& c)
    {
#pragma unroll
        for (int k = 0; k < 6; ++k) { c.s[k] = 0.0f; c.t[k] = 0.0; }
    }
    template <bool kU = false, bool kM = true>
    __device__ static __forceinline__ void interact(const IReg& r, const Slot& s, Acc& c,
                                                    float e2u = float(0))
    {
        float dx = (r.xh - s.v[0]) + (r.xl - s.v[4]);
        float dy = (r.yh - s.v[1]) + (r.yl - s.v[5]);
        float dz = (r.zh - s.v[2]) + (r.zl - s.v[6]);
        float e2 = kU ? e2u : r.e2 + s.v[7];
        float ux = r.vx - s.v[8];
        float uy = r.vy - s.v[9];
        float uz = r.vz - s.v[10];
        float rv = __builtin_fmaf(dz, uz, __builtin_fmaf(dy, uy, dx * ux));
        const float y = split_rsqrt<kRefine, kM || !kU>(dx, dy, dz, e2);
        float q2 = y * y;
        float q3 = q2 * (y * s.v[3]);
        float alpha = 3.0f * rv * q2;
        ux = __builtin_fmaf(-alpha, dx, ux);
        uy = __builtin_fmaf(-alpha, dy, uy);
        uz = __builtin_fmaf(-alpha, dz, uz);
        c.s[0] = __builtin_fmaf(-q3, dx, c.s[0]);
        c.s[1] = __builtin_fmaf(-q3, dy, c.s[1]);
        c.s[2] = __builtin_fmaf(-q3, dz, c.s[2]);
        c.s[3] = __builtin_fmaf(-q3, ux, c.s[3]);
        c.s[4] = __builtin_fmaf(-q3, uy, c.s[4]);
        c.s[5] = __builtin_fmaf(-q3, uz, c.s[5]);
    }
    __device__ static void flush(Acc& c)
    {
#pragma unroll
        for (int k = 0; k < 6; ++k) {
            c.t[k] = __builtin_fma(static_cast<double>(c.s[k]), split_scale<kRefine>(), c.t[k]);
            c.s[k] = 0.0f;
        }
    }
    __device__ static double get(const Acc& c, int k) { return c.t[k]; }
    __device__ static void combine(double* t, const double* p)
    {
#pragma unroll
        for (int k = 0; k < 6; ++k) t[k] += p[k];
    }
    __device__ static void store(const Args& a, int i, const double* t)
    {
#pragma unroll
        for (int k = 0; k < 6; ++k) a.out[k][i] = t[k];
    }
};

// Split-FP32 Tstep: the pair terms of OpTstep in FP32 on split relative
// coordinates (frame as for AccJerk), the sum of w^2 flushed into FP64 every
// kFlush j-particles, the maximum kept in FP32 (it is one pair's term, so it
// carries that term's FP32 error and no accumulated one).  The pair terms
// reach ga ~ m^2 / R^6, beyond FP32 for R below ~1e-6 in N-body units with
// total masses near 1 (a softening of 1e-4 keeps every R above it); an i
// whose FP32 sums overflow there is evaluated again by OpTstep<Fast64>
// (engine.h: fallback64).
struct OpSplitTstep {
    using Args = PairArgs<double>;
    using Frame = SplitFrame;
    using Fallback64 = OpTstep<Fast64>;
    // xh yh zh m | xl yl zl e2 | vx vy vz pad
    using Slot = Slots<float, 12>;
    struct IReg { float xh, yh, zh, xl, yl, zl, e2, m, vx, vy, vz, eta; };
    struct Acc { float sum, mx; double tsum; };
    static constexpr int kNacc = 2;
    static constexpr int kFlush = 64;
    // uniform softening: when every i of the workgroup and every j of a tile
    // carry the same e2, the engine passes the pair value e2 + e2 once
    static constexpr bool kUniformE2 = true;
    __device__ static float e2_of_i(const Args& a, int i) { return static_cast<float>(a.ip[4][i]); }
    __device__ static float slot_e2(const Slot& s) { return s.v[7]; }

    __device__ static Frame frame(const Args& a, int ibase)
    {
        return OpSplitAccJerkT<false>::frame(a, ibase);
    }
    __device__ static void load_i(const Args& a, const Frame& f, int i, IReg& r)
    {
        split2(a.ip[1][i] - f.cx, r.xh, r.xl);
        split2(a.ip[2][i] - f.cy, r.yh, r.yl);
        split2(a.ip[3][i] - f.cz, r.zh, r.zl);
        r.e2 = static_cast<float>(a.ip[4][i]);
        r.m = static_cast<float>(a.ip[0][i]);
        r.vx = static_cast<float>(a.ip[5][i] - f.cvx);
        r.vy = static_cast<float>(a.ip[6][i] - f.cvy);
        r.vz = static_cast<float>(a.ip[7][i] - f.cvz);
        r.eta = static_cast<float>(a.eta);
        r.m *= 2.0f;
    }
    // masses held doubled (exact): the pair's 2 phi = (2 m_i + 2 m_j) / R
    // comes out of one product, and 2 q2 phi = q2 (2 phi) exactly
    __device__ static void load_j(const Args& a, const Frame& f, int j, Slot& s)
    {
        OpSplitAccJerkT<false>::load_j(a, f, j, s);
        s.v[3] *= 2.0f;
    }
    __device__ static void zero(Acc& c)
    {
        c.sum = 0.0f;
        c.mx = 0.0f;
        c.tsum = 0.0;
    }
    template <bool kU = false>
    __device__ static __forceinline__ void interact(const IReg& r, const Slot& s, Acc& c,
                                                    float e2u = float(0))
    {
        float dx = (r.xh - s.v[0]) + (r.xl - s.v[4]);
        float dy = (r.yh - s.v[1]) + (r.yl - s.v[5]);
        float dz = (r.zh - s.v[2]) + (r.zl - s.v[6]);
        float e2 = kU ? e2u : r.e2 + s.v[7];
        float ux = r.vx - s.v[8];
        float uy = r.vy - s.v[9];
        float uz = r.vz - s.v[10];
        float mt = r.m + s.v[3];
        // R^2 with e2 in the first fma, masked as in split_rsqrt (R^2 > e2)
        const float R2 = __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, __builtin_fmaf(dx, dx, e2)));
        float rv = __builtin_fmaf(dz, uz, __builtin_fmaf(dy, uy, dx * ux));
        float v2 = __builtin_fmaf(uz, uz, __builtin_fmaf(uy, uy, ux * ux));
        const bool mask = R2 > e2;
        // no mask on 1/R: a masked pair's term is zeroed at the end
        float q1 = __builtin_amdgcn_rsqf(R2);
        float q2 = q1 * q1;
        float phi2 = mt * q1;  // 2 phi: mt is the doubled mass sum
        float w2 = q2 * (phi2 + v2);
        float ga = q2 * __builtin_fmaf(q2, phi2, w2);
        ga *= r.eta * __builtin_amdgcn_rsqf(w2);
        w2 = __builtin_fmaf(-ga, rv, w2);
        w2 = mask ? w2 : 0.0f;
        c.sum += w2;
        c.mx = __builtin_fmaxf(w2, c.mx);
    }
    __device__ static void flush(Acc& c)
    {
        c.tsum += static_cast<double>(c.sum);
        c.sum = 0.0f;
    }
    __device__ static double get(const Acc& c, int k)
    {
        return k == 0 ? c.tsum : static_cast<double>(c.mx);
    }
    __device__ static void combine(double* t, const double* p)
    {
        OpTstep<Fast64>::combine(t, p);
    }
    __device__ static void store(const Args& a, int i, const double* t)
    {
        OpTstep<Fast64>::store(a, i, t);
    }
};

// ---------------------------------------------------------------------------
// NREG X step (nreg_kernels_common.h:7-61, nreg_kernels.c:5-65): relative
// positions drifted by dt, then sum_j m_j r, the acceleration and
// sum_j m_j / R (times m_i at the end).
// i, j: m rx ry rz e2 vx vy vz; scalar dt; out: mrx mry mrz ax ay az u.
// ---------------------------------------------------------------------------
template <class P>
struct OpNregX {
    using T = typename P::T;
    using Args = PairArgs<T>;
    using Frame = NoFrame;
    using Slot = Slots<T, 8>;
    struct IReg { T m, x, y, z, e2, vx, vy, vz, dt; };
    struct Acc { T d[7]; };
    static constexpr int kNacc = 7;
    static constexpr int kFlush = 1 << 30;

    __device__ static Frame frame(const Args&, int) { return {}; }
    __device__ static void load_i(const Args& a, const Frame&, int i, IReg& r)
    {
        r.m = a.ip[0][i]; r.x = a.ip[1][i]; r.y = a.ip[2][i]; r.z = a.ip[3][i];
        r.e2 = a.ip[4][i]; r.vx = a.ip[5][i]; r.vy = a.ip[6][i]; r.vz = a.ip[7][i];
        r.dt = a.eta;
    }
    __device__ static void load_j(const Args& a, const Frame&, int j, Slot& s)
    {
#pragma unroll
        for (int f = 0; f < 8; ++f) s.v[f] = a.jp[f][j];
    }
    __device__ static void zero(Acc& c)
    {
#pragma unroll
        for (int k = 0; k < 7; ++k) c.d[k] = 0;
    }
    __device__ static __forceinline__ void interact(const IReg& r, const Slot& s, Acc& c)
    {
        T dx = r.x - s.v[1];
        T dy = r.y - s.v[2];
        T dz = r.z - s.v[3];
        const T e2 = r.e2 + s.v[4];
        const T ux = r.vx - s.v[5];
        const T uy = r.vy - s.v[6];
        const T uz = r.vz - s.v[7];
        dx = mad<P>(ux, r.dt, dx);
        dy = mad<P>(uy, r.dt, dy);
        dz = mad<P>(uz, r.dt, dz);
        const T r2 = dot3<P>(dx, dy, dz, dx, dy, dz);
        T q1, q2;
        P::q12(r2, e2, r2 > T(0), q1, q2);
        T q3 = q2 * q1;
        q3 *= s.v[0];
        c.d[0] = mad<P>(s.v[0], dx, c.d[0]);
        c.d[1] = mad<P>(s.v[0], dy, c.d[1]);
        c.d[2] = mad<P>(s.v[0], dz, c.d[2]);
        c.d[3] = nmad<P>(q3, dx, c.d[3]);
        c.d[4] = nmad<P>(q3, dy, c.d[4]);
        c.d[5] = nmad<P>(q3, dz, c.d[5]);
        c.d[6] = mad<P>(s.v[0], q1, c.d[6]);
    }
    __device__ static void flush(Acc&) {}
    __device__ static double get(const Acc& c, int k) { return static_cast<double>(c.d[k]); }
    __device__ static void combine(double* t, const double* p)
    {
#pragma unroll
        for (int k = 0; k < 7; ++k) t[k] += p[k];
    }
    __device__ static void store(const Args& a, int i, const double* t)
    {
#pragma unroll
        for (int k = 0; k < 6; ++k) a.out[k][i] = static_cast<T>(t[k]);
        a.out[6][i] = a.ip[0][i] * static_cast<T>(t[6]);
    }
};

// ---------------------------------------------------------------------------
// NREG V step (nreg_kernels_common.h:63-100, nreg_kernels.c:68-116):
// relative velocities kicked by dt, then sum_j m_j v and sum_j m_j v^2
// (times m_i at the end).
// i, j: m vx vy vz ax ay az; scalar dt; out: mvx mvy mvz k.
// ---------------------------------------------------------------------------
template <class P>
struct OpNregV {
    using T = typename P::T;
    using Args = PairArgs<T>;
    using Frame = NoFrame;
    using Slot = Slots<T, 8>;
    struct IReg { T vx, vy, vz, ax, ay, az, dt; };
    struct Acc { T d[4]; };
    static constexpr int kNacc = 4;
    static constexpr int kFlush = 1 << 30;

    __device__ static Frame frame(const Args&, int) { return {}; }
    __device__ static void load_i(const Args& a, const Frame&, int i, IReg& r)
    {
        r.vx = a.ip[1][i]; r.vy = a.ip[2][i]; r.vz = a.ip[3][i];
        r.ax = a.ip[4][i]; r.ay = a.ip[5][i]; r.az = a.ip[6][i];
        r.dt = a.eta;
    }
    __device__ static void load_j(const Args& a, const Frame&, int j, Slot& s)
    {
#pragma unroll
        for (int f = 0; f < 7; ++f) s.v[f] = a.jp[f][j];
        s.v[7] = 0;
    }
    __device__ static void zero(Acc& c)
    {
#pragma unroll
        for (int k = 0; k < 4; ++k) c.d[k] = 0;
    }
    __device__ static __forceinline__ void interact(const IReg& r, const Slot& s, Acc& c)
    {
        T ux = r.vx - s.v[1];
        T uy = r.vy - s.v[2];
        T uz = r.vz - s.v[3];
        const T bx = r.ax - s.v[4];
        const T by = r.ay - s.v[5];
        const T bz = r.az - s.v[6];
        ux = mad<P>(bx, r.dt, ux);
        uy = mad<P>(by, r.dt, uy);
        uz = mad<P>(bz, r.dt, uz);
        const T v2 = dot3<P>(ux, uy, uz, ux, uy, uz);
        c.d[0] = mad<P>(s.v[0], ux, c.d[0]);
        c.d[1] = mad<P>(s.v[0], uy, c.d[1]);
        c.d[2] = mad<P>(s.v[0], uz, c.d[2]);
        c.d[3] = mad<P>(s.v[0], v2, c.d[3]);
    }
    __device__ static void flush(Acc&) {}
    __device__ static double get(const Acc& c, int k) { return static_cast<double>(c.d[k]); }
    __device__ static void combine(double* t, const double* p)
    {
#pragma unroll
        for (int k = 0; k < 4; ++k) t[k] += p[k];
    }
    __device__ static void store(const Args& a, int i, const double* t)
    {
#pragma unroll
        for (int k = 0; k < 3; ++k) a.out[k][i] = static_cast<T>(t[k]);
        a.out[3][i] = a.ip[0][i] * static_cast<T>(t[3]);
    }
};

// The split path uses the raw v_rsq_f32 estimate (within 1.6 ulp) with its
// mean bias compensated (split_scale): on the 65536-particle Plummer state a
// Newton step leaves the median and the 99.9th percentile of the relative
// force error where they are (other FP32 roundings dominate them) and costs
// four more instructions per pair.  TUPAN_SPLIT_REFINE=1 builds the refined
// form.
#ifndef TUPAN_SPLIT_REFINE
#define TUPAN_SPLIT_REFINE 0
#endif
constexpr bool kSplitRefine = TUPAN_SPLIT_REFINE != 0;
using OpSplitAcc = OpSplitAccT<kSplitRefine>;
using OpSplitAccJerk = OpSplitAccJerkT<kSplitRefine>;

}  // namespace tupan
