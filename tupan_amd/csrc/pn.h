// Post-Newtonian pair acceleration (pnacc_kernel_common.h:8-70,
// pn_terms.h:126-566) as a policy for the tile engine.
//
// The acceleration of i due to j is A r + B v (r = r_i - r_j, v = v_i - v_j)
// with
//   A = (m_j/R^3) (a7 + a6 + a5 + a4 + a2),   B = (m_j/R^2) (b7 + ... + b2),
// where a_k, b_k are the 1/c^k terms of the two-body harmonic-gauge equations
// of motion: 1PN (k = 2), 2PN (4), 2.5PN (5, radiation reaction), 3PN (6)
// and 3.5PN (7).  The run-time `order` selects them cumulatively: 2 and 3
// give 1PN, 4 adds 2PN, 5 adds 2.5PN, 6 adds 3PN, 7 and above add 3.5PN;
// 0 and 1 give zero.  `order` is uniform over the launch, so the level tests
// are scalar branches.
//
// Every polynomial is a chain of mad/nmad in the grouping the reference
// evaluates (left to right within each parenthesis).  Under Ieee64 (no
// contraction) the chain rounds exactly as the reference C code, so one j
// split reproduces it bit for bit; under Fast64/Fast32 the same chain is
// contracted into FMAs.
//
// The reference's 2.5PN b-term carries -v^2 where the published
// harmonic-gauge equations (Blanchet, Living Rev. Relativ. 17 (2014) 2,
// eq. 203) have -4/5 v^2.  The kernel keeps the reference's value.
#pragma once

#include "ops.h"

namespace tupan {

// Invariants of one pair shared by the PN orders.
template <class T>
struct PnPair {
    T im, jm, im2, jm2, imjm;
    T ir, ir2;                       // 1/R, 1/R^2
    T v2, iv2, jv2, iv4, jv4;
    T ivjv, ivjv2, nv, nv2, niv, njv, niv2, njv2, nivnjv;
};

template <class T>
__device__ __forceinline__ constexpr T q(int p, int d)
{
    return static_cast<T>(p) / static_cast<T>(d);
}

template <class T>
__device__ __forceinline__ constexpr T pn_pi2()
{
    return static_cast<T>(9.869604401089358);
}

// 1PN: pn_terms.h:126-151.
template <class P, class T>
__device__ __forceinline__ void pn_1pn(const PnPair<T>& p, T& a, T& b)
{
    T s = nmad<P>(T(2), p.jv2, -p.iv2);
    s = mad<P>(T(4), p.ivjv, s);
    s = mad<P>(q<T>(3, 2), p.njv2, s);
    a = mad<P>(p.ir, mad<P>(T(5), p.im, T(4) * p.jm), s);
    b = nmad<P>(T(3), p.njv, T(4) * p.niv);
}

// 2PN: pn_terms.h:154-219.
template <class P, class T>
__device__ __forceinline__ void pn_2pn(const PnPair<T>& p, T& a, T& b)
{
    T s = T(-2) * (p.jv4 + p.ivjv2);
    s = mad<P>(T(4) * p.jv2, p.ivjv, s);
    T w = q<T>(3, 2) * p.iv2;
    w = mad<P>(q<T>(9, 2), p.jv2, w);
    w = nmad<P>(T(6), p.ivjv, w);
    w = nmad<P>(q<T>(15, 8), p.njv2, w);
    s = mad<P>(p.njv2, w, s);
    w = q<T>(57, 4) * p.im2;
    w = mad<P>(T(9), p.jm2, w);
    w = mad<P>(q<T>(69, 2), p.imjm, w);
    s = nmad<P>(p.ir2, w, s);
    T ci = q<T>(-15, 4) * p.iv2;
    ci = mad<P>(q<T>(5, 4), p.jv2, ci);
    ci = nmad<P>(q<T>(5, 2), p.ivjv, ci);
    ci = mad<P>(q<T>(39, 2), p.niv2, ci);
    ci = nmad<P>(T(39), p.nivnjv, ci);
    ci = mad<P>(q<T>(17, 2), p.njv2, ci);
    T cj = T(4) * p.jv2;
    cj = nmad<P>(T(8), p.ivjv, cj);
    cj = mad<P>(T(2), p.niv2, cj);
    cj = nmad<P>(T(4), p.nivnjv, cj);
    cj = nmad<P>(T(6), p.njv2, cj);
    a = mad<P>(p.ir, mad<P>(p.im, ci, p.jm * cj), s);

    T e = mad<P>(q<T>(55, 4), p.njv, q<T>(-63, 4) * p.niv);
    T u = p.ir * nmad<P>(p.jm * T(2), p.niv + p.njv, p.im * e);
    u = mad<P>(p.iv2, p.njv, u);
    u = nmad<P>(p.ivjv * p.nv, T(4), u);
    u = mad<P>(p.jv2, nmad<P>(T(5), p.njv, T(4) * p.niv), u);
    b = mad<P>(p.njv2, mad<P>(q<T>(9, 2), p.njv, T(-6) * p.niv), u);
}

// 2.5PN: pn_terms.h:222-244, without the common m_i/(R c^5) factor.
template <class P, class T>
__device__ __forceinline__ void pn_25pn(const PnPair<T>& p, T& a, T& b)
{
    a = p.nv * mad<P>(p.ir, mad<P>(q<T>(208, 15), p.jm, q<T>(-24, 5) * p.im),
                      q<T>(12, 5) * p.v2);
    b = mad<P>(p.ir, nmad<P>(q<T>(32, 5), p.jm, q<T>(8, 5) * p.im), -p.v2);
}

// 3PN: pn_terms.h:247-389.
template <class P, class T>
__device__ __forceinline__ void pn_3pn(const PnPair<T>& p, T& a, T& b)
{
    const T niv = p.niv, njv = p.njv, iv2 = p.iv2, jv2 = p.jv2;
    const T ivjv = p.ivjv, njv2 = p.njv2, niv2 = p.niv2;

    // a: eight groups, summed left to right
    T x = T(3) * p.ivjv2;
    x = mad<P>(q<T>(3, 2) * iv2, jv2, x);
    x = nmad<P>(T(12) * ivjv, jv2, x);
    x = mad<P>(q<T>(15, 2), p.jv4, x);
    T y = mad<P>(q<T>(15, 2), nmad<P>(q<T>(1, 4), iv2, ivjv - jv2), q<T>(35, 16) * njv2);
    x = mad<P>(njv2, y, x);
    T s = njv2 * x;
    s = mad<P>(T(2) * jv2, mad<P>(jv2, T(2) * ivjv - jv2, -p.ivjv2), s);

    T g1 = T(244) * ivjv;
    g1 = nmad<P>(q<T>(205, 2), iv2, g1);
    g1 = nmad<P>(q<T>(283, 2), jv2, g1);
    g1 = mad<P>(q<T>(383, 2), njv2, g1);
    T g2 = q<T>(229, 4) * nmad<P>(T(2), ivjv, iv2 + jv2);
    g2 = nmad<P>(q<T>(723, 4), njv2, g2);
    g2 = mad<P>(niv, q<T>(171, 2) * nmad<P>(q<T>(1, 4), niv, njv), g2);
    T t = niv * mad<P>(niv, g2, njv * g1);
    T h = q<T>(191, 4) * iv2;
    h = mad<P>(q<T>(259, 4), jv2, h);
    h = nmad<P>(q<T>(225, 2), ivjv, h);
    h = nmad<P>(q<T>(455, 8), njv2, h);
    t = mad<P>(njv2, h, t);
    h = q<T>(91, 2) * iv2;
    h = mad<P>(T(43), jv2, h);
    h = nmad<P>(q<T>(177, 4), ivjv, h);
    t = mad<P>(ivjv, h, t);
    t = nmad<P>(q<T>(91, 8) * iv2, mad<P>(T(2), jv2, iv2), t);
    t = nmad<P>(q<T>(81, 8), p.jv4, t);
    s = mad<P>(p.im * p.ir, t, s);

    T w = nmad<P>(q<T>(3, 2), niv2, T(3) * (ivjv - jv2));
    w = mad<P>(njv, mad<P>(q<T>(3, 2), njv, T(3) * niv), w);
    t = mad<P>(njv, mad<P>(njv, w, niv * (ivjv - jv2)), p.jv4);
    t = mad<P>(ivjv, nmad<P>(T(2), jv2, ivjv), t);
    s = mad<P>(p.jm * p.ir * T(4), t, s);

    t = mad<P>(T(2), p.nivnjv, -niv2);
    t = mad<P>(q<T>(43, 2), njv2, t);
    t = mad<P>(T(18), ivjv, t);
    t = nmad<P>(T(9), jv2, t);
    s = mad<P>(p.jm2 * p.ir2, t, s);

    t = q<T>(415, 8) * niv2;
    t = nmad<P>(q<T>(375, 4), p.nivnjv, t);
    t = mad<P>(q<T>(1113, 8), njv2, t);
    t = mad<P>(T(18), iv2, t);
    t = mad<P>(pn_pi2<T>(), nmad<P>(q<T>(615, 64), p.nv2, q<T>(123, 64) * p.v2), t);
    t = mad<P>(T(33), nmad<P>(q<T>(1, 2), jv2, ivjv), t);
    s = mad<P>(p.imjm * p.ir2, t, s);

    t = q<T>(-2069, 8) * niv2;
    t = mad<P>(T(543), p.nivnjv, t);
    t = nmad<P>(q<T>(939, 4), njv2, t);
    t = mad<P>(q<T>(471, 8), iv2, t);
    t = mad<P>(q<T>(357, 8), nmad<P>(T(2), ivjv, jv2), t);
    s = mad<P>(p.im2 * p.ir2, t, s);

    constexpr T k41 = q<T>(41, 16) * pn_pi2<T>();
    t = T(16) * p.jm * p.jm2;
    t = mad<P>(p.im2 * p.jm, q<T>(547, 3) - k41, t);
    t = nmad<P>(q<T>(13, 12) * p.im, p.im2, t);
    t = mad<P>(p.im * p.jm2, q<T>(545, 3) - k41, t);
    a = mad<P>(p.ir * p.ir2, t, s);

    // b: five groups
    T r = T(6) * nmad<P>(q<T>(1, 4), iv2, T(2) * jv2 - ivjv);
    r = mad<P>(njv, q<T>(15, 2) * nmad<P>(q<T>(3, 4), njv, niv), r);
    T u = mad<P>(njv, r, T(6) * niv * nmad<P>(T(2), jv2, ivjv));
    T z = jv2 * nmad<P>(T(7), jv2, mad<P>(T(8), ivjv, iv2));
    z = nmad<P>(T(2), p.ivjv2, z);
    z = mad<P>(njv, u, z);
    s = njv * z;
    s = mad<P>(T(4) * niv, nmad<P>(ivjv, jv2, p.jv4), s);

    w = T(4) * nmad<P>(q<T>(1, 2), niv2, ivjv - jv2);
    w = mad<P>(njv, T(2) * mad<P>(T(4), niv, njv), w);
    s = mad<P>(p.jm * p.ir, mad<P>(njv, w, T(2) * niv * (ivjv - jv2)), s);

    T g = q<T>(207, 8) * iv2;
    g = mad<P>(q<T>(81, 8), jv2, g);
    g = nmad<P>(T(36), ivjv, g);
    g = nmad<P>(q<T>(269, 4), njv2, g);
    g = mad<P>(niv, nmad<P>(q<T>(243, 4), niv, q<T>(565, 4) * njv), g);
    h = q<T>(83, 8) * jv2;
    h = mad<P>(q<T>(27, 4), ivjv, h);
    h = nmad<P>(q<T>(137, 8), iv2, h);
    h = nmad<P>(q<T>(95, 12), njv2, h);
    s = mad<P>(p.im * p.ir, mad<P>(njv, h, niv * g), s);

    t = p.jm2 * mad<P>(T(5), njv, T(4) * niv);
    t = mad<P>(p.im2, nmad<P>(q<T>(357, 4), njv, q<T>(311, 4) * niv), t);
    constexpr T k123 = q<T>(123, 32) * pn_pi2<T>();
    t = mad<P>(p.imjm, mad<P>(k123, p.nv, nmad<P>(q<T>(307, 8), niv, q<T>(479, 8) * njv)), t);
    b = mad<P>(p.ir2, t, s);
}

// 3.5PN: pn_terms.h:392-489, without the common m_i/(R c^7) factor.
template <class P, class T>
__device__ __forceinline__ void pn_35pn(const PnPair<T>& p, T& a, T& b)
{
    const T niv = p.niv, njv = p.njv, iv2 = p.iv2, jv2 = p.jv2;
    const T ivjv = p.ivjv, njv2 = p.njv2, niv2 = p.niv2;
    const T nivnjv = p.nivnjv, nv = p.nv, nv2 = p.nv2, v2 = p.v2;

    T s = p.im2 * p.ir2 * nmad<P>(q<T>(4328, 105), njv, q<T>(3992, 105) * niv);
    s = mad<P>(p.imjm * p.ir * p.ir2, mad<P>(q<T>(2872, 21), njv, q<T>(-13576, 105) * niv), s);
    s = mad<P>(p.jm2 * p.ir * p.ir2, q<T>(-3172, 21) * nv, s);

    T g = T(48) * niv2;
    g = nmad<P>(q<T>(4888, 105), iv2, g);
    g = mad<P>(q<T>(2056, 21), ivjv, g);
    g = nmad<P>(q<T>(1028, 21), jv2, g);
    T t = mad<P>(nivnjv, mad<P>(q<T>(744, 5), njv, q<T>(-696, 5) * niv), niv * g);
    T k = q<T>(-288, 5) * njv2;
    k = mad<P>(q<T>(5056, 105), iv2, k);
    k = nmad<P>(q<T>(2224, 21), ivjv, k);
    k = mad<P>(q<T>(5812, 105), jv2, k);
    s = mad<P>(p.im * p.ir, mad<P>(njv, k, t), s);

    g = q<T>(-582, 5) * niv2;
    g = nmad<P>(q<T>(2864, 35), ivjv, g);
    g = mad<P>(q<T>(1432, 35), jv2, g);
    t = mad<P>(nivnjv, nmad<P>(q<T>(1954, 5), njv, q<T>(1746, 5) * niv), niv * g);
    t = mad<P>(q<T>(3568, 105) * nv, iv2, t);
    k = T(158) * njv2;
    k = nmad<P>(q<T>(5752, 105), jv2, k);
    k = mad<P>(q<T>(10048, 105), ivjv, k);
    s = mad<P>(p.jm * p.ir, mad<P>(njv, k, t), s);

    T x = nv * nmad<P>(q<T>(246, 35), p.iv4, T(-56) * nv2 * nv2);
    T y = T(60) * niv2;
    y = nmad<P>(T(180), nivnjv, y);
    y = mad<P>(T(174), njv2, y);
    y = v2 * y;
    y = mad<P>(ivjv, mad<P>(q<T>(984, 35), jv2, q<T>(1068, 35) * (iv2 - ivjv)), y);
    y = nmad<P>(q<T>(534, 35) * iv2, jv2, y);
    y = nmad<P>(q<T>(204, 35), p.jv4, y);
    x = mad<P>(niv, y, x);
    T w = q<T>(-984, 35) * iv2;
    w = mad<P>(q<T>(180, 7), ivjv, w);
    w = nmad<P>(q<T>(732, 35), jv2, w);
    T z = mad<P>(ivjv, w, T(-54) * njv2 * v2);
    z = mad<P>(q<T>(90, 7) * iv2, jv2, z);
    z = mad<P>(q<T>(24, 7), p.jv4, z);
    x = mad<P>(njv, z, x);
    a = s + x;

    s = -(p.im2 * p.ir2 * q<T>(184, 21));
    s = mad<P>(p.imjm * p.ir2, q<T>(6224, 105), s);
    s = mad<P>(p.jm2 * p.ir2, q<T>(6388, 105), s);
    g = q<T>(52, 15) * niv2;
    g = nmad<P>(q<T>(56, 15), nivnjv, g);
    g = nmad<P>(q<T>(44, 15), njv2, g);
    g = nmad<P>(q<T>(132, 35), iv2, g);
    g = mad<P>(q<T>(152, 35), ivjv, g);
    g = nmad<P>(q<T>(48, 35), jv2, g);
    s = mad<P>(p.im * p.ir, g, s);
    g = q<T>(454, 15) * niv2;
    g = nmad<P>(q<T>(372, 5), nivnjv, g);
    g = mad<P>(q<T>(854, 15), njv2, g);
    g = nmad<P>(q<T>(152, 21), iv2, g);
    g = mad<P>(q<T>(2864, 105), ivjv, g);
    g = nmad<P>(q<T>(1768, 105), jv2, g);
    s = mad<P>(p.jm * p.ir, g, s);
    T d = T(60) * nv2 * nv2;
    y = q<T>(-348, 5) * niv2;
    y = mad<P>(q<T>(684, 5), nivnjv, y);
    y = nmad<P>(T(66), njv2, y);
    d = mad<P>(v2, y, d);
    d = mad<P>(q<T>(334, 35), p.iv4, d);
    w = q<T>(-1336, 35) * iv2;
    w = mad<P>(q<T>(1308, 35), ivjv, w);
    w = nmad<P>(q<T>(1252, 35), jv2, w);
    d = mad<P>(ivjv, w, d);
    d = mad<P>(q<T>(654, 35) * iv2, jv2, d);
    d = mad<P>(q<T>(292, 35), p.jv4, d);
    b = s + d;
}

// ---------------------------------------------------------------------------
// PNAcc.  i: m rx ry rz e2 vx vy vz ; j: the same (padded with v_j^2, m_j^2,
// v_j^4 to 12 slots).  PairArgs::order selects the PN level and
// PairArgs::clight holds 1/c^2, 1/c^4, 1/c^5, 1/c^6, 1/c^7.
// ---------------------------------------------------------------------------
template <class P>
struct OpPNAcc {
    using T = typename P::T;
    using Args = PairArgs<T>;
    using Frame = NoFrame;
    using Slot = Slots<T, 12>;
    struct IReg { T m, x, y, z, e2, vx, vy, vz, v2, m2, v4, c2, c4, c5, c6, c7; int order; };
    struct Acc { T ax, ay, az; };
    static constexpr int kNacc = 3;
    static constexpr int kFlush = 1 << 30;

    __device__ static Frame frame(const Args&, int) { return {}; }
    __device__ static void load_i(const Args& a, const Frame&, int i, IReg& r)
    {
        r.m = a.ip[0][i]; r.x = a.ip[1][i]; r.y = a.ip[2][i]; r.z = a.ip[3][i];
        r.e2 = a.ip[4][i]; r.vx = a.ip[5][i]; r.vy = a.ip[6][i]; r.vz = a.ip[7][i];
        r.v2 = dot3<P>(r.vx, r.vy, r.vz, r.vx, r.vy, r.vz);
        r.m2 = r.m * r.m;
        r.v4 = r.v2 * r.v2;
        r.c2 = a.clight[0]; r.c4 = a.clight[1]; r.c5 = a.clight[2];
        r.c6 = a.clight[3]; r.c7 = a.clight[4];
        r.order = a.order;
    }
    __device__ static void load_j(const Args& a, const Frame&, int j, Slot& s)
    {
#pragma unroll
        for (int f = 0; f < 8; ++f) s.v[f] = a.jp[f][j];
        s.v[8] = dot3<P>(s.v[5], s.v[6], s.v[7], s.v[5], s.v[6], s.v[7]);
        s.v[9] = s.v[0] * s.v[0];
        s.v[10] = s.v[8] * s.v[8];
        s.v[11] = 0;
    }
    __device__ static void zero(Acc& c) { c.ax = 0; c.ay = 0; c.az = 0; }
    __device__ static __forceinline__ void interact(const IReg& r, const Slot& s, Acc& c)
    {
        const T rx = r.x - s.v[1];
        const T ry = r.y - s.v[2];
        const T rz = r.z - s.v[3];
        const T e2 = r.e2 + s.v[4];
        const T vx = r.vx - s.v[5];
        const T vy = r.vy - s.v[6];
        const T vz = r.vz - s.v[7];
        const T r2 = dot3<P>(rx, ry, rz, rx, ry, rz);
        PnPair<T> p;
        p.v2 = dot3<P>(vx, vy, vz, vx, vy, vz);
        T ir, ir2;
        P::q12(r2, e2, r2 > T(0), ir, ir2);
        const T ir3 = ir2 * ir;
        const T nx = rx * ir, ny = ry * ir, nz = rz * ir;
        p.im = r.m;
        p.jm = s.v[0];
        p.ir = ir;
        p.ir2 = ir2;
        p.iv2 = r.v2;
        p.jv2 = s.v[8];
        p.niv = dot3<P>(nx, ny, nz, r.vx, r.vy, r.vz);
        p.njv = dot3<P>(nx, ny, nz, s.v[5], s.v[6], s.v[7]);
        p.njv2 = p.njv * p.njv;
        p.ivjv = dot3<P>(r.vx, r.vy, r.vz, s.v[5], s.v[6], s.v[7]);
        T a2, b2;
        pn_1pn<P>(p, a2, b2);
        T sa = a2 * r.c2;
        T sb = b2 * r.c2;
        if (r.order > 3) {
            p.im2 = r.m2;
            p.jm2 = s.v[9];
            p.imjm = r.m * s.v[0];
            p.jv4 = s.v[10];
            p.ivjv2 = p.ivjv * p.ivjv;
            p.nv = dot3<P>(nx, ny, nz, vx, vy, vz);
            p.niv2 = p.niv * p.niv;
            p.nivnjv = p.niv * p.njv;
            T a4, b4;
            pn_2pn<P>(p, a4, b4);
            a4 *= r.c4;
            b4 *= r.c4;
            if (r.order > 4) {
                const T f5 = (r.m * ir) * r.c5;
                T a5, b5;
                pn_25pn<P>(p, a5, b5);
                a5 *= f5;
                b5 *= f5;
                if (r.order > 5) {
                    p.nv2 = p.nv * p.nv;
                    T a6, b6;
                    pn_3pn<P>(p, a6, b6);
                    a6 *= r.c6;
                    b6 *= r.c6;
                    if (r.order > 6) {
                        const T f7 = (r.m * ir) * r.c7;
                        p.iv4 = r.v4;
                        T a7, b7;
                        pn_35pn<P>(p, a7, b7);
                        a6 = mad<P>(a7, f7, a6);
                        b6 = mad<P>(b7, f7, b6);
                    }
                    a5 += a6;
                    b5 += b6;
                }
                a4 += a5;
                b4 += b5;
            }
            sa = mad<P>(a2, r.c2, a4);
            sb = mad<P>(b2, r.c2, b4);
        }
        const T A = sa * (p.jm * ir3);
        const T B = sb * (p.jm * ir2);
        c.ax += mad<P>(A, rx, B * vx);
        c.ay += mad<P>(A, ry, B * vy);
        c.az += mad<P>(A, rz, B * vz);
    }
    __device__ static void flush(Acc&) {}
    __device__ static double get(const Acc& c, int k)
    {
        return static_cast<double>(k == 0 ? c.ax : (k == 1 ? c.ay : c.az));
    }
    __device__ static void combine(double* t, const double* p)
    {
        t[0] += p[0];
        t[1] += p[1];
        t[2] += p[2];
    }
    __device__ static void store(const Args& a, int i, const double* t)
    {
        a.out[0][i] = static_cast<T>(t[0]);
        a.out[1][i] = static_cast<T>(t[1]);
        a.out[2][i] = static_cast<T>(t[2]);
    }
};

// ---------------------------------------------------------------------------
// Split-FP32 PNAcc: FP64 inputs and outputs, the pair terms in FP32.
//
// The separation r is formed from hi + lo FP32 parts of the coordinates
// relative to the workgroup's first i-particle (as the split-FP32 forces do),
// the relative velocity from FP32 velocities relative to that particle, and
// the absolute velocities that the PN terms also read (v_i^2, v_i.v_j, n.v_i,
// ...) are rounded to FP32 directly.  1/R is the v_rsq_f32 estimate; the
// polynomials are the chains above, contracted into FP32 FMAs; the per-pair
// A r + B v is summed in FP32 and flushed into FP64 every kFlush j-particles.
// The PN acceleration is a correction of relative size (v/c)^2 to the
// Newtonian one, so its FP32 relative error (~1e-6) is ~1e-10 of the total
// force for v/c ~ 0.01.  FP32 range: the 3PN terms reach m^4/R^5, beyond
// FP32 for R below ~1e-6 in N-body units; an i whose FP32 sums overflow is
// evaluated again by OpPNAcc<Fast64> after the call (engine.h:
// fallback_kernel).
// ---------------------------------------------------------------------------
struct OpSplitPNAcc {
    using Args = PairArgs<double>;
    using Frame = SplitFrame;
    // sums that overflow FP32 are recomputed in FP64 after the call
    // (engine.h: fallback_kernel)
    using PostFallback64 = OpPNAcc<Fast64>;
    using P = Fast32;
    // xh yh zh m | xl yl zl e2 | ux uy uz v2 | vx vy vz m2 | v4 pad pad pad
    using Slot = Slots<float, 20>;
    struct IReg {
        float xh, yh, zh, xl, yl, zl, e2, m, ux, uy, uz, vx, vy, vz, v2, m2, v4;
        float c2, c4, c5, c6, c7;
        int order;
    };
    struct Acc { float s[3]; double t[3]; };
    static constexpr int kNacc = 3;
    static constexpr int kFlush = 64;

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
        r.m = static_cast<float>(a.ip[0][i]);
        r.ux = static_cast<float>(a.ip[5][i] - f.cvx);
        r.uy = static_cast<float>(a.ip[6][i] - f.cvy);
        r.uz = static_cast<float>(a.ip[7][i] - f.cvz);
        r.vx = static_cast<float>(a.ip[5][i]);
        r.vy = static_cast<float>(a.ip[6][i]);
        r.vz = static_cast<float>(a.ip[7][i]);
        r.v2 = dot3<P>(r.vx, r.vy, r.vz, r.vx, r.vy, r.vz);
        r.m2 = r.m * r.m;
        r.v4 = r.v2 * r.v2;
        r.c2 = static_cast<float>(a.clight[0]);
        r.c4 = static_cast<float>(a.clight[1]);
        r.c5 = static_cast<float>(a.clight[2]);
        r.c6 = static_cast<float>(a.clight[3]);
        r.c7 = static_cast<float>(a.clight[4]);
        r.order = a.order;
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
        s.v[12] = static_cast<float>(a.jp[5][j]);
        s.v[13] = static_cast<float>(a.jp[6][j]);
        s.v[14] = static_cast<float>(a.jp[7][j]);
        s.v[11] = dot3<P>(s.v[12], s.v[13], s.v[14], s.v[12], s.v[13], s.v[14]);
        s.v[15] = s.v[3] * s.v[3];
        s.v[16] = s.v[11] * s.v[11];
        s.v[17] = s.v[18] = s.v[19] = 0.0f;
    }
    __device__ static void zero(Acc& c)
    {
#pragma unroll
        for (int k = 0; k < 3; ++k) { c.s[k] = 0.0f; c.t[k] = 0.0; }
    }
    __device__ static __forceinline__ void interact(const IReg& r, const Slot& s, Acc& c)
    {
        const float rx = (r.xh - s.v[0]) + (r.xl - s.v[4]);
        const float ry = (r.yh - s.v[1]) + (r.yl - s.v[5]);
        const float rz = (r.zh - s.v[2]) + (r.zl - s.v[6]);
        const float e2 = r.e2 + s.v[7];
        const float vx = r.ux - s.v[8];
        const float vy = r.uy - s.v[9];
        const float vz = r.uz - s.v[10];
        const float r2 = dot3<P>(rx, ry, rz, rx, ry, rz);
        float ir = __builtin_amdgcn_rsqf(r2 + e2);
        ir = r2 > 0.0f ? ir : 0.0f;
        const float ir2 = ir * ir;
        const float nx = rx * ir, ny = ry * ir, nz = rz * ir;
        PnPair<float> p;
        p.v2 = dot3<P>(vx, vy, vz, vx, vy, vz);
        p.im = r.m;
        p.jm = s.v[3];
        p.ir = ir;
        p.ir2 = ir2;
        p.iv2 = r.v2;
        p.jv2 = s.v[11];
        p.niv = dot3<P>(nx, ny, nz, r.vx, r.vy, r.vz);
        p.njv = dot3<P>(nx, ny, nz, s.v[12], s.v[13], s.v[14]);
        p.njv2 = p.njv * p.njv;
        p.ivjv = dot3<P>(r.vx, r.vy, r.vz, s.v[12], s.v[13], s.v[14]);
        float a2, b2;
        pn_1pn<P>(p, a2, b2);
        float sa = a2 * r.c2;
        float sb = b2 * r.c2;
        if (r.order > 3) {
            p.im2 = r.m2;
            p.jm2 = s.v[15];
            p.imjm = r.m * s.v[3];
            p.jv4 = s.v[16];
            p.ivjv2 = p.ivjv * p.ivjv;
            p.nv = dot3<P>(nx, ny, nz, vx, vy, vz);
            p.niv2 = p.niv * p.niv;
            p.nivnjv = p.niv * p.njv;
            float a4, b4;
            pn_2pn<P>(p, a4, b4);
            a4 *= r.c4;
            b4 *= r.c4;
            if (r.order > 4) {
                const float f5 = (r.m * ir) * r.c5;
                float a5, b5;
                pn_25pn<P>(p, a5, b5);
                a5 *= f5;
                b5 *= f5;
                if (r.order > 5) {
                    p.nv2 = p.nv * p.nv;
                    float a6, b6;
                    pn_3pn<P>(p, a6, b6);
                    a6 *= r.c6;
                    b6 *= r.c6;
                    if (r.order > 6) {
                        const float f7 = (r.m * ir) * r.c7;
                        p.iv4 = r.v4;
                        float a7, b7;
                        pn_35pn<P>(p, a7, b7);
                        a6 = __builtin_fmaf(a7, f7, a6);
                        b6 = __builtin_fmaf(b7, f7, b6);
                    }
                    a5 += a6;
                    b5 += b6;
                }
                a4 += a5;
                b4 += b5;
            }
            sa = __builtin_fmaf(a2, r.c2, a4);
            sb = __builtin_fmaf(b2, r.c2, b4);
        }
        const float A = sa * (p.jm * (ir2 * ir));
        const float B = sb * (p.jm * ir2);
        c.s[0] += __builtin_fmaf(A, rx, B * vx);
        c.s[1] += __builtin_fmaf(A, ry, B * vy);
        c.s[2] += __builtin_fmaf(A, rz, B * vz);
    }
    __device__ static void flush(Acc& c)
    {
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            c.t[k] += static_cast<double>(c.s[k]);
            c.s[k] = 0.0f;
        }
    }
    __device__ static double get(const Acc& c, int k) { return c.t[k]; }
    __device__ static void combine(double* t, const double* p)
    {
        t[0] += p[0]; t[1] += p[1]; t[2] += p[2];
    }
    __device__ static void store(const Args& a, int i, const double* t)
    {
        a.out[0][i] = t[0];
        a.out[1][i] = t[1];
        a.out[2][i] = t[2];
    }
};

}  // namespace tupan
