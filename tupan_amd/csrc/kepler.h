// Device universal-variable Kepler solver (FP64), following
// tupan/lib/src/universal_kepler_solver.h:17-614 and the Sakura pair drift of
// sakura_kernel_common.h:45-243.
//
// Differences from the reference, none of which changes a converged value:
//   - The reference evaluates six Stumpff functions per Laguerre step, each
//     with its own sqrt and cos/cosh or sin/sinh of the same argument
//     (f, f', f'' at one s, :201-222).  Here the four c_k share one sqrt and
//     one sincos (or sinh, cosh) per step; the values are the same numbers.
//   - The two step-subdivision retry loops (:486-499, :580-606) are
//     unbounded in the reference.  Here each is capped at kMaxDoublings
//     halvings; a pair that reaches the cap keeps its last state and raises
//     the status flag, which the host turns into an error.
//   - Work counters (solves, Laguerre steps, subdivisions, leapfrogs).
// The file is compiled with -ffp-contract=off so every expression keeps the
// reference's operation order; transcendentals come from the device math
// library and may differ from the host libm in the last bit.
#pragma once

#include "common.h"

namespace tupan {
namespace kep {

constexpr double kTol = 2.2737367544323205948e-13;  // 2^-42
constexpr int kMaxIter = 64;
constexpr int kMaxDoublings = 20;
// Solves one pair may spend across both retry loops (enough for the energy
// check to double its step count up to 2^kMaxDoublings); past it every further
// solve reports failure at once, so the loops run out within a few hundred
// trivial steps and the pair is flagged.
constexpr unsigned int kSolveBudget = 1u << 22;
constexpr double kPi = 3.141592653589793;

struct Counters {
    unsigned int budget = kSolveBudget;
    unsigned int solves = 0;
    unsigned int iters = 0;
    unsigned int retries = 0;
    unsigned int leapfrog = 0;
    unsigned int fallbacks = 0;
    int fail = 0;
};

__device__ __forceinline__ int sgn(double x) { return (x > 0) - (x < 0); }

// S_k(s, alpha) = s^k c_k(alpha s^2), k = 0..3 (:17-120).
struct Stumpff {
    double S0, S1, S2, S3;
};

__device__ __forceinline__ Stumpff stumpff(double s, double alpha)
{
    const double s2 = s * s;
    const double s3 = s * s2;
    const double z = alpha * s2;
    double c0, c1, c2, c3;
    if (z < 0) {
        const double t = sqrt(-z);
        double sn, cs;
        sincos(t, &sn, &cs);
        c0 = cs;
        c1 = sn / t;
        c2 = (cs - 1) / z;
        c3 = (sn / t - 1) / z;
    } else if (z > 0) {
        const double t = sqrt(z);
        const double ch = cosh(t);
        const double sh = sinh(t);
        c0 = ch;
        c1 = sh / t;
        c2 = (ch - 1) / z;
        c3 = (sh / t - 1) / z;
    } else {
        c0 = 1.0;
        c1 = 1.0;
        c2 = 1 / 2.0;
        c3 = 1 / 6.0;
    }
    Stumpff r;
    r.S0 = c0;
    r.S1 = s * c1;
    r.S2 = s2 * c2;
    r.S3 = s3 * c3;
    return r;
}

// Laguerre iteration of order 5 on Kepler's equation in s (:225-254).
// k = {dt, r0, r0.v0, m, alpha}.  Returns 0, -1 (h == 0), -2 (MAXITER),
// -3 (sign of s changed).
__device__ inline int laguerre5(double x0, double& x, const double* k, Counters& cnt,
                                int max_iter)
{
    int it = 0;
    double delta;
    x = x0;
    const double mar = k[3] + k[4] * k[1];
    do {
        const Stumpff S = stumpff(x, k[4]);
        const double fv = (k[1] * x + k[2] * S.S2 + mar * S.S3) - k[0];
        const double dfv = k[1] + k[2] * S.S1 + mar * S.S2;
        const double ddfv = k[2] * S.S0 + mar * S.S1;
        const double a = dfv;
        const double a2 = a * a;
        const double b = a2 - fv * ddfv;
        const double g = 5 * fv;
        const double h = a + sgn(a) * sqrt(fabs(4 * (5 * b - a2)));
        cnt.iters++;
        if (h == 0) return -1;
        delta = -g / h;
        x += delta;
        it += 1;
        if (it > max_iter) return -2;
    } while (fabs(delta) > kTol);
    if (sgn(x) != sgn(x0)) return -3;
    return 0;
}

// Lagrange f, g and their derivatives applied to (p, v) (:312-357).
__device__ inline void lagrange_update(double dt, double s, double r0, double m, double e2,
                                       double alpha, double* p, double* v)
{
    const Stumpff S = stumpff(s, alpha);
    const double x0 = p[0], y0 = p[1], z0 = p[2];
    const double u0 = v[0], w0 = v[1], t0 = v[2];
    const double lf = 1 - m * S.S2 / r0;
    const double lg = dt - m * S.S3;
    const double x1 = x0 * lf + u0 * lg;
    const double y1 = y0 * lf + w0 * lg;
    const double z1 = z0 * lf + t0 * lg;
    const double r1sq = x1 * x1 + y1 * y1 + z1 * z1;
    const double r1 = sqrt(r1sq + e2);
    const double ldf = -m * S.S1 / (r0 * r1);
    const double ldg = (1 + lg * ldf) / lf;
    p[0] = x1;
    p[1] = y1;
    p[2] = z1;
    v[0] = x0 * ldf + u0 * ldg;
    v[1] = y0 * ldf + w0 * ldg;
    v[2] = z0 * ldf + t0 * ldg;
}

// Dimensionless description of one solve, the inputs of a learned initial
// guess: with r the (softened) separation, tau = dt sqrt(m / r^3) (dt after
// the period reduction), sigma = r.v / sqrt(m r), kappa = v^2 r / m,
// eps = e2 / (r^2 + e2); the target is delta = s r / dt - 1 for the converged
// universal anomaly s (0 when s = dt / r, the plain guess).
struct SolveFeatures {
    float tau, sigma, kappa, eps;
};

// Learned guess for one solve: use = 1 replaces the plain s0 by
// (dt / r) (1 + delta); the Laguerre iteration then gets kLearnedIter steps
// before it restarts from the plain s0 (fallback counted in cnt.fallbacks).
struct Guess {
    int use = 0;
    double delta = 0.0;
};
constexpr int kLearnedIter = 8;

__device__ __forceinline__ bool solve_setup(double dt0, double m, double e2, const double* p,
                                            const double* v, double k[5], double& s0,
                                            SolveFeatures* f)
{
    double r2 = p[0] * p[0] + p[1] * p[1] + p[2] * p[2];
    if (!(r2 > 0)) return false;
    const double rr2 = r2;
    r2 += e2;
    const double r = sqrt(r2);
    const double v2 = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
    const double rv = p[0] * v[0] + p[1] * v[1] + p[2] * v[2];
    const double beta = 2 - (e2 / r2);
    const double alpha = v2 - beta * m / r;

    double dt = dt0;
    if (alpha < 0) {
        const double a = m / fabs(alpha);
        const double T = 2 * kPi * a * sqrt(a / m);
        const double ratio = dt0 / T;
        dt = (ratio - static_cast<double>(static_cast<long>(ratio))) * T;
    }

    s0 = dt / r;
    if (alpha > 0) {
        // Bate, Mueller & White (1971) eq. 4.5.11 guess for hyperbolae
        const double sal = sqrt(alpha);
        const double ss = fabs(2 * alpha * dt / (rv + (m + alpha * r) / sal));
        if (ss > 1) s0 = sgn(dt) * log(ss) / sal;
    }
    k[0] = dt; k[1] = r; k[2] = rv; k[3] = m; k[4] = alpha;
    if (f) {
        f->tau = static_cast<float>(dt * sqrt(m / (r2 * r)));
        f->sigma = static_cast<float>(rv / sqrt(m * r));
        f->kappa = static_cast<float>(v2 * r / m);
        f->eps = static_cast<float>(e2 / (rr2 + e2));
    }
    return true;
}

// One solve over dt0 (:360-463).  p, v in/out.  With a learned guess the
// iteration starts from it and falls back to the plain s0 after
// kLearnedIter steps without convergence (or on any solver error).
__device__ inline int kepler_once(double dt0, double m, double e2, double* p, double* v,
                                  Counters& cnt, const Guess* guess = nullptr,
                                  double* s_out = nullptr)
{
    double k[5], s0;
    if (!solve_setup(dt0, m, e2, p, v, k, s0, nullptr)) return 0;
    if (cnt.budget == 0) return -4;
    cnt.budget--;
    cnt.solves++;
    double s;
    int err = 1;
    if (guess && guess->use) {
        const double sg = (k[0] / k[1]) * (1 + guess->delta);
        if (sgn(sg) == sgn(s0)) err = laguerre5(sg, s, k, cnt, kLearnedIter);
        if (err != 0) cnt.fallbacks++;
    }
    if (err != 0) err = laguerre5(s0, s, k, cnt, kMaxIter);
    if (err == 0) {
        if (s_out) *s_out = s;
        lagrange_update(k[0], s, k[1], m, e2, k[4], p, v);
    }
    return err;
}

// Retry with n = 2, 4, ... equal substeps on any solver error (:466-507).
__device__ inline void kepler_subdivided(double dt, double m, double e2, const double* p0,
                                         const double* v0, double* p, double* v, Counters& cnt,
                                         const Guess* guess = nullptr, double* s_out = nullptr)
{
    for (int lvl = 0; lvl <= kMaxDoublings; ++lvl) {
        const long n = 1L << lvl;
        bool ok = true;
        p[0] = p0[0]; p[1] = p0[1]; p[2] = p0[2];
        v[0] = v0[0]; v[1] = v0[1]; v[2] = v0[2];
        for (long i = 0; i < n; ++i) {
            // only the undivided first solve uses the learned guess, and
            // only it reports its converged anomaly
            const bool first = lvl == 0;
            if (kepler_once(dt / n, m, e2, p, v, cnt, first ? guess : nullptr,
                            first ? s_out : nullptr) != 0) {
                ok = false;
                break;
            }
        }
        if (ok) return;
        cnt.retries++;
        if (cnt.budget == 0) break;
    }
    cnt.fail = 1;
}

// Full solver with the softened-energy check and step doubling (:510-614).
__device__ inline void kepler_solve(double dt, double m, double e2, const double* p0,
                                    const double* v0, double* p1, double* v1, Counters& cnt,
                                    const Guess* guess = nullptr, double* s_out = nullptr)
{
    cnt.budget = kSolveBudget;
    double p[3], v[3];
    kepler_subdivided(dt, m, e2, p0, v0, p, v, cnt, guess, s_out);
    if (e2 != 0) {
        double r2 = p0[0] * p0[0] + p0[1] * p0[1] + p0[2] * p0[2];
        if (!(r2 > 0)) {
            p[0] = p0[0]; p[1] = p0[1]; p[2] = p0[2];
            v[0] = v0[0]; v[1] = v0[1]; v[2] = v0[2];
        } else {
            r2 += e2;
            double r = sqrt(r2);
            double v2 = v0[0] * v0[0] + v0[1] * v0[1] + v0[2] * v0[2];
            const double u0 = 2 * m / r;
            const double en0 = v2 - u0;

            r2 = p[0] * p[0] + p[1] * p[1] + p[2] * p[2];
            r2 += e2;
            r = sqrt(r2);
            v2 = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
            double u1 = 2 * m / r;
            double en1 = v2 - u1;

            const double tol = 64 * kTol;
            double err = fabs(en1 - en0);
            if (!(2 * err < tol * (u1 + u0))) {
                long n = 1;
                int lvl;
                for (lvl = 1; lvl <= kMaxDoublings; ++lvl) {
                    n *= 2;
                    bool ok = true;
                    p[0] = p0[0]; p[1] = p0[1]; p[2] = p0[2];
                    v[0] = v0[0]; v[1] = v0[1]; v[2] = v0[2];
                    for (long i = 0; i < n; ++i) {
                        const double q[3] = {p[0], p[1], p[2]};
                        const double w[3] = {v[0], v[1], v[2]};
                        kepler_subdivided(dt / n, m, e2, q, w, p, v, cnt);
                        r2 = p[0] * p[0] + p[1] * p[1] + p[2] * p[2];
                        r2 += e2;
                        r = sqrt(r2);
                        v2 = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
                        u1 = 2 * m / r;
                        en1 = v2 - u1;
                        err = fabs(en1 - en0);
                        if (2 * err > tol * (u1 + u0)) {
                            ok = false;
                            break;
                        }
                    }
                    if (ok) break;
                    cnt.retries++;
                }
                if (lvl > kMaxDoublings) cnt.fail = 1;
                if (cnt.budget == 0) cnt.fail = 1;
            }
        }
    }
    p1[0] = p[0]; p1[1] = p[1]; p1[2] = p[2];
    v1[0] = v[0]; v1[1] = v[1]; v1[2] = v[2];
}

// Leapfrog DKD with one softened kick, for far pairs (:45-91).
__device__ inline void leapfrog(double dt, double m, double e2, double* p, double* v,
                                Counters& cnt)
{
    const double h = dt / 2;
    double x = p[0], y = p[1], z = p[2];
    double u = v[0], w = v[1], t = v[2];
    x += u * h;
    y += w * h;
    z += t * h;
    const double r2 = x * x + y * y + z * z;
    double q2 = 1 / (r2 + e2);
    q2 = r2 > 0 ? q2 : 0.0;
    const double q3 = q2 * sqrt(q2);
    const double mq = m * q3;
    u += (-mq * x) * dt;
    w += (-mq * y) * dt;
    t += (-mq * z) * dt;
    x += u * h;
    y += w * h;
    z += t * h;
    p[0] = x; p[1] = y; p[2] = z;
    v[0] = u; v[1] = w; v[2] = t;
    cnt.leapfrog++;
}

// Far pairs (r > 64 m / v^2) take one leapfrog step (:94-123).
__device__ inline void twobody(double dt, double m, double e2, double* p, double* v,
                               Counters& cnt, const Guess* guess = nullptr,
                               double* s_out = nullptr)
{
    const double r2 = p[0] * p[0] + p[1] * p[1] + p[2] * p[2];
    const double v2 = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
    const double R = 64 * (m / v2);
    if (r2 > R * R) {
        leapfrog(dt, m, e2, p, v, cnt);
    } else {
        const double p0[3] = {p[0], p[1], p[2]};
        const double v0[3] = {v[0], v[1], v[2]};
        kepler_solve(dt, m, e2, p0, v0, p, v, cnt, guess, s_out);
    }
}

// evolve_twobody flags (:126-191); any other flag leaves the pair unchanged.
// A learned guess (and the anomaly report) serves the first twobody stage.
__device__ inline void drift_twobody(double dt, long flag, double m, double e2, double* p,
                                     double* v, Counters& cnt, const Guess* guess = nullptr,
                                     double* s_out = nullptr)
{
    if (flag == -1) {
        p[0] -= v[0] * dt; p[1] -= v[1] * dt; p[2] -= v[2] * dt;
        twobody(dt, m, e2, p, v, cnt, guess, s_out);
    } else if (flag == 1) {
        twobody(dt, m, e2, p, v, cnt, guess, s_out);
        p[0] -= v[0] * dt; p[1] -= v[1] * dt; p[2] -= v[2] * dt;
    } else if (flag == -2) {
        p[0] -= v[0] * dt / 2; p[1] -= v[1] * dt / 2; p[2] -= v[2] * dt / 2;
        twobody(dt, m, e2, p, v, cnt, guess, s_out);
        p[0] -= v[0] * dt / 2; p[1] -= v[1] * dt / 2; p[2] -= v[2] * dt / 2;
    } else if (flag == 2) {
        twobody(dt / 2, m, e2, p, v, cnt, guess, s_out);
        p[0] -= v[0] * dt; p[1] -= v[1] * dt; p[2] -= v[2] * dt;
        twobody(dt / 2, m, e2, p, v, cnt);
    }
}

// The reference's DKD leapfrog with 1/R^3 from the refined hardware
// reciprocal square root (within an ulp of the IEEE division and square
// root of leapfrog()).  Not used by the Sakura kernels: the increment
// p1 - p0 of a far pair cancels its O(v dt) terms, so last-ulp changes in
// 1/R^3 show up amplified by v / (a dt) in the summed drifts.
__device__ inline void leapfrog_fast(double dt, double m, double e2, double* p, double* v,
                                     Counters& cnt)
{
    const double h = dt / 2;
    double x = p[0], y = p[1], z = p[2];
    double u = v[0], w = v[1], t = v[2];
    x += u * h;
    y += w * h;
    z += t * h;
    const double r2 = x * x + y * y + z * z;
    double q1 = rsqrt_refined(r2 + e2);
    q1 = r2 > 0 ? q1 : 0.0;
    const double mq = m * (q1 * q1 * q1);
    u += (-mq * x) * dt;
    w += (-mq * y) * dt;
    t += (-mq * z) * dt;
    x += u * h;
    y += w * h;
    z += t * h;
    p[0] = x; p[1] = y; p[2] = z;
    v[0] = u; v[1] = w; v[2] = t;
    cnt.leapfrog++;
}

// Exact two-body advance of particles i and j in place, in their
// centre-of-mass frame (kepler_solver_kernel_common.h:8-81): the SIA n = 2
// leaves.
__device__ inline void advance_pair(int i, int j, const double* m, double* rx, double* ry,
                                    double* rz, const double* e2, double* vx, double* vy,
                                    double* vz, double dt, Counters& cnt)
{
    const double r0[3] = {rx[i] - rx[j], ry[i] - ry[j], rz[i] - rz[j]};
    const double ee = e2[i] + e2[j];
    const double v0[3] = {vx[i] - vx[j], vy[i] - vy[j], vz[i] - vz[j]};
    const double mt = m[i] + m[j];
    const double mu0 = m[i] / mt;
    const double mu1 = m[j] / mt;
    double cx = mu0 * rx[i] + mu1 * rx[j];
    double cy = mu0 * ry[i] + mu1 * ry[j];
    double cz = mu0 * rz[i] + mu1 * rz[j];
    const double wx = mu0 * vx[i] + mu1 * vx[j];
    const double wy = mu0 * vy[i] + mu1 * vy[j];
    const double wz = mu0 * vz[i] + mu1 * vz[j];
    cx += wx * dt;
    cy += wy * dt;
    cz += wz * dt;
    double pr[3], pv[3];
    kepler_solve(dt, mt, ee, r0, v0, pr, pv, cnt);
    rx[i] = cx + mu1 * pr[0];
    ry[i] = cy + mu1 * pr[1];
    rz[i] = cz + mu1 * pr[2];
    vx[i] = wx + mu1 * pv[0];
    vy[i] = wy + mu1 * pv[1];
    vz[i] = wz + mu1 * pv[2];
    rx[j] = cx - mu0 * pr[0];
    ry[j] = cy - mu0 * pr[1];
    rz[j] = cz - mu0 * pr[2];
    vx[j] = wx - mu0 * pv[0];
    vy[j] = wy - mu0 * pv[1];
    vz[j] = wz - mu0 * pv[2];
}

// Far-pair part of drift_twobody: runs the flag's stages as long as every
// twobody stage is a far pair (leapfrog) and returns true with p, v
// advanced; returns false (p, v partly advanced, to be discarded) at the
// first stage that needs a Kepler solve.  The leapfrogs of a pair that
// returns false are not counted.  The far/near test is the reference's
// (IEEE division), so a single-stage flag classifies every pair exactly as
// drift_twobody does.
template <bool kFast>
__device__ inline bool far_stage(double dt, double m, double e2, double* p, double* v,
                                 Counters& cnt)
{
    const double r2 = p[0] * p[0] + p[1] * p[1] + p[2] * p[2];
    const double v2 = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
    const double R = 64 * (m / v2);
    if (!(r2 > R * R)) return false;
    if (kFast) {
        leapfrog_fast(dt, m, e2, p, v, cnt);
    } else {
        leapfrog(dt, m, e2, p, v, cnt);
    }
    return true;
}

// drift_twobody_far for a flag known at compile time (kFlag: -1, 1, -2, 2;
// any other value leaves the pair unchanged).  The same operations in the
// same order as the run-time form below, without the flag tests in the
// pair loop.
template <long kFlag>
__device__ __forceinline__ bool drift_twobody_far_t(double dt, double m, double e2, double* p,
                                                    double* v, Counters& cnt)
{
    Counters c;
    bool ok = true;
    if constexpr (kFlag == -1) {
        p[0] -= v[0] * dt; p[1] -= v[1] * dt; p[2] -= v[2] * dt;
        ok = far_stage<false>(dt, m, e2, p, v, c);
    } else if constexpr (kFlag == 1) {
        ok = far_stage<false>(dt, m, e2, p, v, c);
        if (ok) { p[0] -= v[0] * dt; p[1] -= v[1] * dt; p[2] -= v[2] * dt; }
    } else if constexpr (kFlag == -2) {
        p[0] -= v[0] * dt / 2; p[1] -= v[1] * dt / 2; p[2] -= v[2] * dt / 2;
        ok = far_stage<false>(dt, m, e2, p, v, c);
        if (ok) { p[0] -= v[0] * dt / 2; p[1] -= v[1] * dt / 2; p[2] -= v[2] * dt / 2; }
    } else if constexpr (kFlag == 2) {
        ok = far_stage<false>(dt / 2, m, e2, p, v, c);
        if (ok) {
            p[0] -= v[0] * dt; p[1] -= v[1] * dt; p[2] -= v[2] * dt;
            ok = far_stage<false>(dt / 2, m, e2, p, v, c);
        }
    }
    if (ok) cnt.leapfrog += c.leapfrog;
    return ok;
}

template <bool kFast = false>
__device__ inline bool drift_twobody_far(double dt, long flag, double m, double e2, double* p,
                                         double* v, Counters& cnt)
{
    Counters c;
    bool ok = true;
    if (flag == -1) {
        p[0] -= v[0] * dt; p[1] -= v[1] * dt; p[2] -= v[2] * dt;
        ok = far_stage<kFast>(dt, m, e2, p, v, c);
    } else if (flag == 1) {
        ok = far_stage<kFast>(dt, m, e2, p, v, c);
        if (ok) { p[0] -= v[0] * dt; p[1] -= v[1] * dt; p[2] -= v[2] * dt; }
    } else if (flag == -2) {
        p[0] -= v[0] * dt / 2; p[1] -= v[1] * dt / 2; p[2] -= v[2] * dt / 2;
        ok = far_stage<kFast>(dt, m, e2, p, v, c);
        if (ok) { p[0] -= v[0] * dt / 2; p[1] -= v[1] * dt / 2; p[2] -= v[2] * dt / 2; }
    } else if (flag == 2) {
        ok = far_stage<kFast>(dt / 2, m, e2, p, v, c);
        if (ok) {
            p[0] -= v[0] * dt; p[1] -= v[1] * dt; p[2] -= v[2] * dt;
            ok = far_stage<kFast>(dt / 2, m, e2, p, v, c);
        }
    }
    if (ok) cnt.leapfrog += c.leapfrog;
    return ok;
}

// 1/x from the hardware estimate refined by two Newton steps (within an ulp
// of IEEE division for normal x).
__device__ __forceinline__ double recip_refined(double x)
{
    double y = __builtin_amdgcn_rcp(x);
    double e = __builtin_fma(-x, y, 1.0);
    y = __builtin_fma(y, e, y);
    e = __builtin_fma(-x, y, 1.0);
    return __builtin_fma(y, e, y);
}

// State of a pair's first twobody stage (the one a learned guess serves):
// the relative position after the flag's pre-drift and that stage's dt.
__device__ inline double first_stage(double dt, long flag, double* p, const double* v)
{
    if (flag == -1) {
        p[0] -= v[0] * dt; p[1] -= v[1] * dt; p[2] -= v[2] * dt;
        return dt;
    }
    if (flag == -2) {
        p[0] -= v[0] * dt / 2; p[1] -= v[1] * dt / 2; p[2] -= v[2] * dt / 2;
        return dt;
    }
    if (flag == 2) return dt / 2;
    return dt;
}

}  // namespace kep
}  // namespace tupan
