// Sakura pair drift and batched two-body Kepler advance for gfx950.
//
// sakura: for every (i, j) pair the relative orbit is advanced by
// drift_twobody (sakura_kernel_common.h:126-191) and the mass-weighted
// changes are summed per i (:194-243, sakura_kernel.c:5-63).  It runs on the
// LDS tile engine with one i per thread; lanes diverge between the leapfrog
// and Kepler branches and over Laguerre iteration counts.
//
// kepler_pairs: one thread per independent particle pair, the centre-of-
// mass frame two-body advance of kepler_solver_kernel_common.h:8-81, written
// back in place (the SIA n = 2 leaves, integrator/fewbody.py:30-42).
//
// Compiled with -ffp-contract=off (reference operation order).
#include "engine.h"
#include "kepler.h"
#include "tupan_internal.h"

namespace tupan {

struct SakuraArgs {
    const double* ip[8];
    const double* jp[8];
    double* out[6];
    int ni, nj;
    int nsplit, jchunk;
    double* partial;
    double dt;
    long flag;
    unsigned long long* counters;
    int* status;
};

// Wave-level sum so that one lane per wave issues the counter atomics.
__device__ __forceinline__ unsigned int wave_sum(unsigned int v)
{
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
    return v;
}

struct OpSakura {
    using Args = SakuraArgs;
    struct Frame {};
    template <class T, int N>
    struct alignas(16) S {
        T v[N];
    };
    using Slot = S<double, 8>;
    struct IReg { double m, x, y, z, e2, vx, vy, vz, dt; long flag; };
    struct Acc { double d[6]; kep::Counters c; };
    static constexpr int kNacc = 6;
    static constexpr int kFlush = 1 << 30;

    __device__ static Frame frame(const Args&, int) { return {}; }
    __device__ static void load_i(const Args& a, const Frame&, int i, IReg& r)
    {
        r.m = a.ip[0][i]; r.x = a.ip[1][i]; r.y = a.ip[2][i]; r.z = a.ip[3][i];
        r.e2 = a.ip[4][i]; r.vx = a.ip[5][i]; r.vy = a.ip[6][i]; r.vz = a.ip[7][i];
        r.dt = a.dt;
        r.flag = a.flag;
    }
    __device__ static void load_j(const Args& a, const Frame&, int j, Slot& s)
    {
#pragma unroll
        for (int f = 0; f < 8; ++f) s.v[f] = a.jp[f][j];
    }
    __device__ static void zero(Acc& c)
    {
#pragma unroll
        for (int k = 0; k < 6; ++k) c.d[k] = 0.0;
        c.c = kep::Counters();
    }
    __device__ static void interact(const IReg& r, const Slot& s, Acc& c)
    {
        const double p0[3] = {r.x - s.v[1], r.y - s.v[2], r.z - s.v[3]};
        const double e2 = r.e2 + s.v[4];
        const double v0[3] = {r.vx - s.v[5], r.vy - s.v[6], r.vz - s.v[7]};
        const double m = r.m + s.v[0];
        double p[3] = {p0[0], p0[1], p0[2]};
        double v[3] = {v0[0], v0[1], v0[2]};
        kep::drift_twobody(r.dt, r.flag, m, e2, p, v, c.c);
        const double mu = s.v[0] / m;
        c.d[0] += mu * (p[0] - p0[0]);
        c.d[1] += mu * (p[1] - p0[1]);
        c.d[2] += mu * (p[2] - p0[2]);
        c.d[3] += mu * (v[0] - v0[0]);
        c.d[4] += mu * (v[1] - v0[1]);
        c.d[5] += mu * (v[2] - v0[2]);
    }
    __device__ static void flush(Acc&) {}
    __device__ static double get(const Acc& c, int k) { return c.d[k]; }
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

// The engine (engine.h) is specialised here with a counter flush: each
// thread's counts are summed over the wave and lane 0 adds them to the
// device counters with vector-memory atomics.
template <int BLOCK, int TILE>
__global__ __launch_bounds__(BLOCK) void sakura_kernel_dev(SakuraArgs a)
{
    using Op = OpSakura;
    __shared__ Op::Slot tile[TILE];
    const GridPos g = grid_pos(a.nsplit);
    const int tid = static_cast<int>(threadIdx.x);
    const int i = g.iblock * BLOCK + tid;
    const bool valid = i < a.ni;
    Op::IReg ir;
    Op::Acc acc;
    Op::load_i(a, {}, valid ? i : a.ni - 1, ir);
    Op::zero(acc);
    const int j0 = g.jsplit * a.jchunk;
    const int j1 = min(a.nj, j0 + a.jchunk);
    for (int jt = j0; jt < j1; jt += TILE) {
        const int cnt = min(TILE, j1 - jt);
        __syncthreads();
        for (int k = tid; k < cnt; k += BLOCK) Op::load_j(a, {}, jt + k, tile[k]);
        __syncthreads();
        if (valid) {
            for (int k = 0; k < cnt; ++k) Op::interact(ir, tile[k], acc);
        }
    }
    if (valid) {
        if (a.nsplit == 1) {
            Op::store(a, i, acc.d);
        } else {
            double* part = a.partial + static_cast<size_t>(g.jsplit) * 6 * a.ni;
#pragma unroll
            for (int k = 0; k < 6; ++k) part[static_cast<size_t>(k) * a.ni + i] = acc.d[k];
        }
    }
    if (a.counters) {
        const unsigned int s = wave_sum(valid ? acc.c.solves : 0u);
        const unsigned int it = wave_sum(valid ? acc.c.iters : 0u);
        const unsigned int rt = wave_sum(valid ? acc.c.retries : 0u);
        const unsigned int lf = wave_sum(valid ? acc.c.leapfrog : 0u);
        if ((tid & (kWave - 1)) == 0) {
            atomicAdd(a.counters + 0, static_cast<unsigned long long>(s));
            atomicAdd(a.counters + 1, static_cast<unsigned long long>(it));
            atomicAdd(a.counters + 2, static_cast<unsigned long long>(rt));
            atomicAdd(a.counters + 3, static_cast<unsigned long long>(lf));
        }
    }
    if (a.status && valid && acc.c.fail) atomicOr(a.status, 1);
}

int sakura_call(size_t ni, const double* const* ip, size_t nj, const double* const* jp,
                double dt, long flag, double* const* out, unsigned long long* counters,
                int* status, void* stream)
{
    if (ni > static_cast<size_t>(INT32_MAX) || nj > static_cast<size_t>(INT32_MAX))
        return set_error(TUPAN_ERR_ARG, "particle count exceeds 2^31-1");
    if (ni == 0) return TUPAN_OK;
    SakuraArgs a{};
    for (int f = 0; f < 8; ++f) {
        if (!ip[f] || (nj > 0 && !jp[f])) return set_error(TUPAN_ERR_ARG, "null input array");
        a.ip[f] = ip[f];
        a.jp[f] = jp[f];
    }
    for (int k = 0; k < 6; ++k) {
        if (!out[k]) return set_error(TUPAN_ERR_ARG, "null output array");
        a.out[k] = out[k];
    }
    a.ni = static_cast<int>(ni);
    a.nj = static_cast<int>(nj);
    a.dt = dt;
    a.flag = flag;
    a.counters = counters;
    a.status = status;
    constexpr int BLOCK = 128, TILE = 64;
    const Device& dev = current_device();
    Launch L = plan_launch(a.ni, a.nj, BLOCK, 1, TILE, 2 * dev.target_wg, 64);
    a.nsplit = L.nsplit;
    a.jchunk = L.jchunk;
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (L.nsplit > 1) {
        a.partial = static_cast<double*>(
            workspace(static_cast<size_t>(L.nsplit) * 6 * ni * sizeof(double), s));
        if (!a.partial) return set_error(TUPAN_ERR_ALLOC, "workspace allocation failed");
    }
    const unsigned grid = static_cast<unsigned>(L.nblock_i) * static_cast<unsigned>(L.nsplit);
    timing_begin(s);
    hipLaunchKernelGGL((sakura_kernel_dev<BLOCK, TILE>), dim3(grid), dim3(BLOCK), 0, s, a);
    timing_end(s);
    if (L.nsplit > 1) {
        const unsigned rg = static_cast<unsigned>((a.ni + 255) / 256);
        hipLaunchKernelGGL((reduce_kernel<OpSakura>), dim3(rg), dim3(256), 0, s, a);
    }
    return check_launch("sakura");
}

// ---------------------------------------------------------------------------
// Independent two-body pairs, in place.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(64) void kepler_pairs_kernel(int npairs, const int* idx,
                                                          const double* m, double* rx,
                                                          double* ry, double* rz,
                                                          const double* e2, double* vx,
                                                          double* vy, double* vz, double dt,
                                                          int* status)
{
    const int p = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    if (p >= npairs) return;
    const int i = idx[2 * p];
    const int j = idx[2 * p + 1];
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
    kep::Counters cnt;
    kep::kepler_solve(dt, mt, ee, r0, v0, pr, pv, cnt);
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
    if (status && cnt.fail) atomicOr(status, 1);
}

int kepler_pairs_call(size_t npairs, const int* idx, const double* m, double* rx, double* ry,
                      double* rz, const double* e2, double* vx, double* vy, double* vz,
                      double dt, int* status, void* stream)
{
    if (npairs == 0) return TUPAN_OK;
    if (npairs > static_cast<size_t>(INT32_MAX / 2))
        return set_error(TUPAN_ERR_ARG, "too many pairs");
    if (!idx || !m || !rx || !ry || !rz || !e2 || !vx || !vy || !vz)
        return set_error(TUPAN_ERR_ARG, "null array");
    const unsigned grid = static_cast<unsigned>((npairs + 63) / 64);
    hipLaunchKernelGGL(kepler_pairs_kernel, dim3(grid), dim3(64), 0,
                       static_cast<hipStream_t>(stream), static_cast<int>(npairs), idx, m, rx,
                       ry, rz, e2, vx, vy, vz, dt, status);
    return check_launch("kepler_pairs");
}

}  // namespace tupan
