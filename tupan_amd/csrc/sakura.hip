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
#include "mlp_wave.h"
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
    // two-pass form (null in the one-pass form): per (j-split, i) the count
    // of deferred near pairs and their j indices (kNearSlots slots each)
    int* near_count;
    int* near_j;
    // when set, dt is read from device memory (device-driven steps)
    const double* dt_dev;
};

__device__ __forceinline__ double sakura_dt(const SakuraArgs& a)
{
    return a.dt_dev ? *a.dt_dev : a.dt;
}

// Near pairs listed per (j-split, i); further ones are solved by
// near_overflow_kernel.
constexpr int kNearSlots = 32;
// Calls with at least this many pairs take the two-pass form by default.
constexpr double kTwoPassPairs = 4194304.0;
// i-particles per thread in the far-pair pass (independent dependency
// chains per thread).
#ifndef TUPAN_SAKURA_FAR_WPT
#define TUPAN_SAKURA_FAR_WPT 1
#endif
constexpr int kFarWpt = TUPAN_SAKURA_FAR_WPT;
// Workgroups the Sakura launch plan aims for, in units of the device's
// target (four 256-thread workgroups per CU).
#ifndef TUPAN_SAKURA_TARGET
#define TUPAN_SAKURA_TARGET 4
#endif
constexpr int kSakuraTarget = TUPAN_SAKURA_TARGET;

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
        r.dt = sakura_dt(a);
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

// Counter flush shared by the kernels: wave sums, one atomic per wave and
// nonzero counter.
__device__ __forceinline__ void flush_counters(unsigned long long* counters, bool valid,
                                               const kep::Counters& c, int tid)
{
    if (!counters) return;
    const unsigned int s = wave_sum(valid ? c.solves : 0u);
    const unsigned int it = wave_sum(valid ? c.iters : 0u);
    const unsigned int rt = wave_sum(valid ? c.retries : 0u);
    const unsigned int lf = wave_sum(valid ? c.leapfrog : 0u);
    if ((tid & (kWave - 1)) == 0) {
        if (s) atomicAdd(counters + 0, static_cast<unsigned long long>(s));
        if (it) atomicAdd(counters + 1, static_cast<unsigned long long>(it));
        if (rt) atomicAdd(counters + 2, static_cast<unsigned long long>(rt));
        if (lf) atomicAdd(counters + 3, static_cast<unsigned long long>(lf));
    }
}

// Counter bank of the near pass.  Its thousands of short waves all finish
// within a few microseconds, and their atomics on the caller's five
// counters serialise in the cache (≈200 µs for 4 570 waves); each workgroup
// adds into one of kBankSlots 128-byte slots instead, and
// near_bank_flush_kernel sums the slots into the counters.  Slot entries:
// solves, Laguerre steps, retries, leapfrog pairs, guess fallbacks.
constexpr int kBankSlots = 64;
constexpr int kBankStride = 16;  // unsigned long longs per slot

__global__ void near_bank_flush_kernel(const unsigned long long* bank,
                                       unsigned long long* counters,
                                       unsigned long long* fallbacks)
{
    const int k = static_cast<int>(threadIdx.x);
    if (k >= 5) return;
    unsigned long long v = 0;
    for (int b = 0; b < kBankSlots; ++b) v += bank[b * kBankStride + k];
    if (v == 0) return;
    if (k < 4) {
        if (counters) atomicAdd(counters + k, v);
    } else if (fallbacks) {
        atomicAdd(fallbacks, v);
    }
}

// ---------------------------------------------------------------------------
// Two-pass form.  Lanes of one wave hold 64 different i against the same j,
// so in the one-pass kernel a single near pair in a wave makes the other 63
// lanes wait through its Kepler solve.  Pass 1 runs the tile loop with far
// pairs only (leapfrog stages) and records the near pairs; the near pairs
// are then compacted into one dense list and solved one per thread (pass 2);
// pass 3 adds each (j-split, i)'s near contributions, in j order, to its far
// partial sum, and the split reduction finishes as in the one-pass form.
// Per (split, i) the far pairs are summed first, then the near pairs, each
// in j order: deterministic.  Each pair's drift is the one-pass arithmetic
// (the leapfrog keeps IEEE division and square root: a pair's increment
// p1 - p0 cancels its O(v dt) terms down to O(a dt^2), so it amplifies any
// change in their rounding by v / (a dt), up to 1e8 here); only the far
// pairs' mass ratio m_j / m comes from the refined hardware reciprocal, a
// factor on the increment, within an ulp of the division.
// ---------------------------------------------------------------------------
template <int BLOCK, int TILE, int WPT, long FLAG>
__global__ __launch_bounds__(BLOCK) void sakura_far_kernel(SakuraArgs a)
{
    using Op = OpSakura;
    __shared__ Op::Slot tile[TILE];
    const GridPos g = grid_pos(a.nsplit);
    const int tid = static_cast<int>(threadIdx.x);
    const int ibase = g.iblock * (BLOCK * WPT);
    Op::IReg r[WPT];
    Op::Acc acc[WPT];
    int* slots[WPT];
    int ndef[WPT];
    bool valid[WPT];
#pragma unroll
    for (int w = 0; w < WPT; ++w) {
        const int i = ibase + w * BLOCK + tid;
        valid[w] = i < a.ni;
        Op::load_i(a, {}, valid[w] ? i : a.ni - 1, r[w]);
        Op::zero(acc[w]);
        const size_t entry = static_cast<size_t>(g.jsplit) * a.ni + (valid[w] ? i : 0);
        slots[w] = a.near_j + entry * kNearSlots;
        ndef[w] = 0;
    }
    const int j0 = g.jsplit * a.jchunk;
    const int j1 = min(a.nj, j0 + a.jchunk);
    for (int jt = j0; jt < j1; jt += TILE) {
        const int cnt = min(TILE, j1 - jt);
        __syncthreads();
        for (int k = tid; k < cnt; k += BLOCK) Op::load_j(a, {}, jt + k, tile[k]);
        __syncthreads();
        for (int k = 0; k < cnt; ++k) {
            const Op::Slot sj = tile[k];
#pragma unroll
            for (int w = 0; w < WPT; ++w) {
                if (!valid[w]) continue;
                const double p0[3] = {r[w].x - sj.v[1], r[w].y - sj.v[2], r[w].z - sj.v[3]};
                const double e2 = r[w].e2 + sj.v[4];
                const double v0[3] = {r[w].vx - sj.v[5], r[w].vy - sj.v[6], r[w].vz - sj.v[7]};
                const double m = r[w].m + sj.v[0];
                double p[3] = {p0[0], p0[1], p0[2]};
                double v[3] = {v0[0], v0[1], v0[2]};
                if (!kep::drift_twobody_far_t<FLAG>(r[w].dt, m, e2, p, v, acc[w].c)) {
                    // near pair: the first kNearSlots go to the list, the
                    // rest to near_overflow_kernel
                    if (ndef[w] < kNearSlots) slots[w][ndef[w]] = jt + k;
                    ++ndef[w];
                    continue;
                }
                const double mu = sj.v[0] * kep::recip_refined(m);
                acc[w].d[0] += mu * (p[0] - p0[0]);
                acc[w].d[1] += mu * (p[1] - p0[1]);
                acc[w].d[2] += mu * (p[2] - p0[2]);
                acc[w].d[3] += mu * (v[0] - v0[0]);
                acc[w].d[4] += mu * (v[1] - v0[1]);
                acc[w].d[5] += mu * (v[2] - v0[2]);
            }
        }
    }
    kep::Counters tot;
#pragma unroll
    for (int w = 0; w < WPT; ++w) {
        const int i = ibase + w * BLOCK + tid;
        if (valid[w]) {
            double* part = a.partial + static_cast<size_t>(g.jsplit) * 6 * a.ni;
#pragma unroll
            for (int k = 0; k < 6; ++k) part[static_cast<size_t>(k) * a.ni + i] = acc[w].d[k];
            a.near_count[static_cast<size_t>(g.jsplit) * a.ni + i] = ndef[w];
            tot.leapfrog += acc[w].c.leapfrog;
            tot.fail |= acc[w].c.fail;
        }
    }
    flush_counters(a.counters, true, tot, tid);
    if (a.status && tot.fail) atomicOr(a.status, 1);
}

// Exclusive scan of the listed counts min(count, kNearSlots) in three
// kernels: per-block sums (kScanPer entries per block, each thread four
// consecutive ones), a one-workgroup scan of the block sums, and the
// per-block scan that writes the offsets.  total[0] receives the listed
// sum, total[1] the number of entries with more near pairs than slots, and
// those entries are appended to `over_list` (in no particular order; each
// is processed independently).
constexpr int kScanThreads = 256;
constexpr int kOverflowWaves = 256;  // grid of the graph-safe overflow pass
constexpr int kScanPer = 4 * kScanThreads;

__device__ __forceinline__ int block_exclusive_scan(int v, int* lds, int& block_total)
{
    const int t = static_cast<int>(threadIdx.x);
    lds[t] = v;
    __syncthreads();
    for (int d = 1; d < kScanThreads; d <<= 1) {
        const int w = t >= d ? lds[t - d] : 0;
        __syncthreads();
        lds[t] += w;
        __syncthreads();
    }
    const int incl = lds[t];
    block_total = lds[kScanThreads - 1];
    __syncthreads();
    return incl - v;
}

__global__ __launch_bounds__(kScanThreads) void near_block_sums_kernel(int n, const int* count,
                                                                       int* bsum, int* bover)
{
    __shared__ int lds[kScanThreads];
    const int base = static_cast<int>(blockIdx.x) * kScanPer + 4 * static_cast<int>(threadIdx.x);
    int sum = 0, ov = 0;
    for (int q = 0; q < 4; ++q) {
        const int e = base + q;
        if (e < n) {
            const int c = count[e];
            sum += min(c, kNearSlots);
            ov += c > kNearSlots;
        }
    }
    int tot_s, tot_o;
    block_exclusive_scan(sum, lds, tot_s);
    block_exclusive_scan(ov, lds, tot_o);
    if (threadIdx.x == 0) {
        bsum[blockIdx.x] = tot_s;
        bover[blockIdx.x] = tot_o;
    }
}

__global__ __launch_bounds__(kScanThreads) void near_block_offsets_kernel(int nb, int* bsum,
                                                                          const int* bover,
                                                                          int* total,
                                                                          unsigned long long* bank)
{
    __shared__ int lds[kScanThreads];
    if (bank) {
        // the near pass's counter bank starts from zero
        for (int k = static_cast<int>(threadIdx.x); k < kBankSlots * kBankStride;
             k += kScanThreads)
            bank[k] = 0;
    }
    int carry = 0, over = 0;
    for (int b0 = 0; b0 < nb; b0 += kScanThreads) {
        const int b = b0 + static_cast<int>(threadIdx.x);
        const int v = b < nb ? bsum[b] : 0;
        const int o = b < nb ? bover[b] : 0;
        int tot, tot_o;
        const int ex = block_exclusive_scan(v, lds, tot);
        block_exclusive_scan(o, lds, tot_o);
        if (b < nb) bsum[b] = carry + ex;
        carry += tot;
        over += tot_o;
    }
    if (threadIdx.x == 0) {
        total[0] = carry;
        total[1] = over;
        total[2] = 0;  // overflow list fill counter
    }
}

__global__ __launch_bounds__(kScanThreads) void near_offsets_kernel(int n, const int* count,
                                                                    const int* boff, int* off,
                                                                    int* total, int* over_list)
{
    __shared__ int lds[kScanThreads];
    const int base = static_cast<int>(blockIdx.x) * kScanPer + 4 * static_cast<int>(threadIdx.x);
    int c[4];
    int sum = 0;
    for (int q = 0; q < 4; ++q) {
        const int e = base + q;
        c[q] = e < n ? count[e] : 0;
        sum += min(c[q], kNearSlots);
    }
    int tot;
    int run = boff[blockIdx.x] + block_exclusive_scan(sum, lds, tot);
    for (int q = 0; q < 4; ++q) {
        const int e = base + q;
        if (e >= n) break;
        off[e] = run;
        run += min(c[q], kNearSlots);
        if (c[q] > kNearSlots) over_list[atomicAdd(total + 2, 1)] = e;
    }
}

// Near-pair list: item = (i, j), grouped by (split, i), j ascending.
// Items at or beyond `cap` are not written (a call whose list is longer
// than its capacity is flagged by near_capacity_kernel).
__global__ __launch_bounds__(256) void near_compact_kernel(int n, int ni, const int* count,
                                                           const int* off, const int* near_j,
                                                           int2* items, int cap)
{
    const int e = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    if (e >= n) return;
    const int c = min(count[e], kNearSlots);
    const int o = off[e];
    const int i = e % ni;
    for (int k = 0; k < c && o + k < cap; ++k)
        items[o + k] = make_int2(i, near_j[static_cast<size_t>(e) * kNearSlots + k]);
}

// Capacity check of the graph-safe form: status bit 2 when the near-pair
// list is longer than the space reserved for it.
__global__ void near_capacity_kernel(const int* total, int cap, int* status)
{
    if (threadIdx.x == 0 && total[0] > cap && status) atomicOr(status, 2);
}

// One near pair, item t of the list: the pair's drift (from the learned
// guess when `guess` is set and usable), its six contributions (contrib is
// SoA [6][nitems]) and, with a log, the first stage's features, the learned
// target and the pair's Laguerre steps.
__device__ __forceinline__ void near_pair(const SakuraArgs& a, int t, int nitems,
                                          const int2* items, double* contrib, bool guess,
                                          float d, const SolveLogDev& log, kep::Counters& c)
{
    const int2 it = items[t];
    const int i = it.x, j = it.y;
    const double p0[3] = {a.ip[1][i] - a.jp[1][j], a.ip[2][i] - a.jp[2][j],
                          a.ip[3][i] - a.jp[3][j]};
    const double e2 = a.ip[4][i] + a.jp[4][j];
    const double v0[3] = {a.ip[5][i] - a.jp[5][j], a.ip[6][i] - a.jp[6][j],
                          a.ip[7][i] - a.jp[7][j]};
    const double mj = a.jp[0][j];
    const double m = a.ip[0][i] + mj;
    double p[3] = {p0[0], p0[1], p0[2]};
    double v[3] = {v0[0], v0[1], v0[2]};
    kep::Guess gs;
    if (guess) {
        gs.use = d == d && d > -0.9f && d < 8.0f;
        gs.delta = static_cast<double>(d);
    }
    const bool logging = log.target != nullptr && t < log.capacity;
    double s_conv = __builtin_nan("");
    kep::drift_twobody(sakura_dt(a), a.flag, m, e2, p, v, c, guess ? &gs : nullptr,
                       logging ? &s_conv : nullptr);
    const double mu = mj / m;
    const size_t n = static_cast<size_t>(nitems);
    contrib[0 * n + t] = mu * (p[0] - p0[0]);
    contrib[1 * n + t] = mu * (p[1] - p0[1]);
    contrib[2 * n + t] = mu * (p[2] - p0[2]);
    contrib[3 * n + t] = mu * (v[0] - v0[0]);
    contrib[4 * n + t] = mu * (v[1] - v0[1]);
    contrib[5 * n + t] = mu * (v[2] - v0[2]);
    if (logging) {
        double q[3] = {p0[0], p0[1], p0[2]};
        const double dts = kep::first_stage(sakura_dt(a), a.flag, q, v0);
        double k[5], s0;
        kep::SolveFeatures f{};
        float target = __builtin_nanf("");
        if (kep::solve_setup(dts, m, e2, q, v0, k, s0, &f) && s_conv == s_conv)
            target = static_cast<float>(s_conv * k[1] / k[0] - 1);
        const size_t cap = static_cast<size_t>(log.capacity);
        log.features[0 * cap + t] = f.tau;
        log.features[1 * cap + t] = f.sigma;
        log.features[2 * cap + t] = f.kappa;
        log.features[3 * cap + t] = f.eps;
        log.target[t] = target;
        log.iterations[t] = static_cast<int>(c.iters);
    }
}

// The near pass's counters (into the workgroup's slot of the counter bank,
// when counting) and failure status.
__device__ __forceinline__ void near_flush(const SakuraArgs& a, bool valid,
                                           const kep::Counters& c, unsigned long long* bank)
{
    if (bank) {
        unsigned long long* slot = bank + (blockIdx.x % kBankSlots) * kBankStride;
        flush_counters(slot, valid, c, static_cast<int>(threadIdx.x));
        const unsigned int fb = wave_sum(valid ? c.fallbacks : 0u);
        if ((threadIdx.x & (kWave - 1)) == 0 && fb)
            atomicAdd(slot + 4, static_cast<unsigned long long>(fb));
    }
    if (a.status && valid && c.fail) atomicOr(a.status, 1);
}

// Pass 2: one near pair per thread.  With `delta` the first solve of the
// pair starts from the learned guess (kepler.h: Guess).
__global__ __launch_bounds__(128) void sakura_near_kernel(SakuraArgs a, int nitems,
                                                          const int* nitems_dev,
                                                          const int2* items, double* contrib,
                                                          const float* delta, SolveLogDev log,
                                                          unsigned long long* bank)
{
    // nitems: the list's capacity (contrib's stride); nitems_dev, when
    // given, the number of items listed on the device
    const int t = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    const int count = nitems_dev ? min(*nitems_dev, nitems) : nitems;
    const bool valid = t < count;
    kep::Counters c;
    if (valid)
        near_pair(a, t, nitems, items, contrib, delta != nullptr, delta ? delta[t] : 0.f, log,
                  c);
    near_flush(a, valid, c, bank);
}

// Pass 2 with the learned guess evaluated in the same kernel: each wave
// forms its 64 pairs' first-stage features (kepler.h: SolveFeatures; NaN for
// a pair whose first stage has no solve setup), runs the network on them
// with the stand-alone inference kernel's arithmetic (mlp_wave.h), then
// solves its pairs from the guesses.  The packed network is staged in LDS.
__global__ __launch_bounds__(128, 2) void sakura_near_learned_kernel(SakuraArgs a, int nitems,
                                                                  const int2* items,
                                                                  double* contrib,
                                                                  const NetPacked* net,
                                                                  SolveLogDev log,
                                                                  unsigned long long* bank)
{
    __shared__ __align__(16) NetPacked lnet;
    mlp_stage(net, &lnet);
    const int t = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    const int lane = static_cast<int>(threadIdx.x) & (kWave - 1);
    const bool valid = t < nitems;
    float xs[4] = {0.f, 0.f, 0.f, 0.f};
    if (valid) {
        const int2 it = items[t];
        const int i = it.x, j = it.y;
        double q[3] = {a.ip[1][i] - a.jp[1][j], a.ip[2][i] - a.jp[2][j],
                       a.ip[3][i] - a.jp[3][j]};
        const double e2 = a.ip[4][i] + a.jp[4][j];
        const double v0[3] = {a.ip[5][i] - a.jp[5][j], a.ip[6][i] - a.jp[6][j],
                              a.ip[7][i] - a.jp[7][j]};
        const double m = a.ip[0][i] + a.jp[0][j];
        const double dts = kep::first_stage(sakura_dt(a), a.flag, q, v0);
        double k[5], s0;
        kep::SolveFeatures f;
        if (!kep::solve_setup(dts, m, e2, q, v0, k, s0, &f))
            f.tau = f.sigma = f.kappa = f.eps = __builtin_nanf("");
        prepare_features(f.tau, f.sigma, f.kappa, f.eps, lnet.mean, lnet.inv_std, xs);
    }
    // lanes l and l + 32 supply the B operands of items l mod 32 and 32 + l mod 32
    float x[2][4];
#pragma unroll
    for (int c = 0; c < 2; ++c) {
#pragma unroll
        for (int q = 0; q < 4; ++q) x[c][q] = __shfl(xs[q], 32 * c + (lane & 31), kWave);
    }
    float y[2];
    mlp_wave_eval(&lnet, x, lane, y);
    const float d = mlp_delta(&lnet, lane < 32 ? y[0] : y[1]);
    // the solve reloads the pair rather than keep the feature pass's values
    // live across the network (register pressure)
    __asm__ volatile("" ::: "memory");
    kep::Counters c;
    if (valid) near_pair(a, t, nitems, items, contrib, true, d, log, c);
    near_flush(a, valid, c, bank);
}

// Pass 3: add each (split, i)'s near contributions, in list order, to its
// far partial sum.
__global__ __launch_bounds__(256) void near_add_kernel(int n, int ni, int nitems,
                                                       const int* count, const int* off,
                                                       const double* contrib, double* partial)
{
    const int e = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    if (e >= n) return;
    const int c = min(count[e], kNearSlots);
    if (c == 0) return;
    const int o = off[e];
    const int split = e / ni;
    const int i = e % ni;
    double* part = partial + static_cast<size_t>(split) * 6 * ni;
    for (int k = 0; k < 6; ++k) {
        double sum = part[static_cast<size_t>(k) * ni + i];
        for (int q = 0; q < c && o + q < nitems; ++q)
            sum += contrib[static_cast<size_t>(k) * nitems + o + q];
        part[static_cast<size_t>(k) * ni + i] = sum;
    }
}

// Entries with more near pairs than list slots, one wave per entry: the
// wave walks its j range in chunks of 64 (one j per lane), ranks the near
// pairs of the chunk in j order with a ballot, solves those ranked after the
// first kNearSlots (in parallel, one per lane), and adds the chunk's
// contributions to the partial sum through a fixed shuffle tree
// (deterministic).
__device__ __forceinline__ void overflow_entry(const SakuraArgs& a, int e, int lane,
                                               kep::Counters& c)
{
    const int split = e / a.ni;
    const int i = e % a.ni;
    const int j0 = split * a.jchunk;
    const int j1 = min(a.nj, j0 + a.jchunk);
    double d[6] = {0, 0, 0, 0, 0, 0};
    int seen = 0;  // near pairs before this chunk
    for (int jb = j0; jb < j1; jb += kWave) {
        const int j = jb + lane;
        const bool inr = j < j1;
        const int jj = inr ? j : j0;
        const double p0[3] = {a.ip[1][i] - a.jp[1][jj], a.ip[2][i] - a.jp[2][jj],
                              a.ip[3][i] - a.jp[3][jj]};
        const double e2 = a.ip[4][i] + a.jp[4][jj];
        const double v0[3] = {a.ip[5][i] - a.jp[5][jj], a.ip[6][i] - a.jp[6][jj],
                              a.ip[7][i] - a.jp[7][jj]};
        const double mj = a.jp[0][jj];
        const double m = a.ip[0][i] + mj;
        double p[3] = {p0[0], p0[1], p0[2]};
        double v[3] = {v0[0], v0[1], v0[2]};
        kep::Counters scratch;
        const bool near = inr && !kep::drift_twobody_far(sakura_dt(a), a.flag, m, e2, p, v, scratch);
        const unsigned long long mask = __ballot(near);
        const int rank = seen + __popcll(mask & ((1ull << lane) - 1ull));
        seen += __popcll(mask);
        double cd[6] = {0, 0, 0, 0, 0, 0};
        if (near && rank >= kNearSlots) {
            p[0] = p0[0]; p[1] = p0[1]; p[2] = p0[2];
            v[0] = v0[0]; v[1] = v0[1]; v[2] = v0[2];
            kep::drift_twobody(sakura_dt(a), a.flag, m, e2, p, v, c);
            const double mu = mj / m;
            cd[0] = mu * (p[0] - p0[0]);
            cd[1] = mu * (p[1] - p0[1]);
            cd[2] = mu * (p[2] - p0[2]);
            cd[3] = mu * (v[0] - v0[0]);
            cd[4] = mu * (v[1] - v0[1]);
            cd[5] = mu * (v[2] - v0[2]);
        }
#pragma unroll
        for (int k = 0; k < 6; ++k) {
            double x = cd[k];
#pragma unroll
            for (int off = kWave / 2; off > 0; off >>= 1) x += __shfl_xor(x, off, kWave);
            d[k] += x;
        }
    }
    if (lane == 0) {
        double* part = a.partial + static_cast<size_t>(split) * 6 * a.ni;
        for (int k = 0; k < 6; ++k) part[static_cast<size_t>(k) * a.ni + i] += d[k];
    }
}

// The entries are walked grid-stride; nover_dev, when given, holds their
// number on the device.
__global__ __launch_bounds__(64) void near_overflow_kernel(SakuraArgs a, const int* over_list,
                                                           int nover, const int* nover_dev)
{
    const int lane = static_cast<int>(threadIdx.x);
    const int nent = nover_dev ? *nover_dev : nover;
    kep::Counters c;
    for (int w = static_cast<int>(blockIdx.x); w < nent; w += static_cast<int>(gridDim.x)) {
        overflow_entry(a, over_list[w], lane, c);
    }
    flush_counters(a.counters, true, c, lane);
    if (a.status && c.fail) atomicOr(a.status, 1);
}

int sakura_call(size_t ni, const double* const* ip, size_t nj, const double* const* jp,
                double dt, long flag, double* const* out, unsigned long long* counters,
                int* status, void* stream)
{
    SakuraOptions opt;
    return sakura_call_ex(ni, ip, nj, jp, dt, flag, out, opt, counters, status, stream);
}

int sakura_call_ex(size_t ni, const double* const* ip, size_t nj, const double* const* jp,
                   double dt, long flag, double* const* out, SakuraOptions& opt,
                   unsigned long long* counters, int* status, void* stream)
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
    a.dt_dev = opt.dt_dev;
    constexpr int BLOCK = 128, TILE = 64;
    const Device& dev = current_device();
    hipStream_t s = static_cast<hipStream_t>(stream);
    bool two_pass = opt.path == 2 ||
                    (opt.path == 0 && static_cast<double>(ni) * static_cast<double>(nj) >=
                                          static_cast<double>(kTwoPassPairs));
    if (opt.net || opt.log.target) two_pass = true;
    const bool graph = opt.graph_capacity >= 0;
    if (graph) {
        if (opt.net || opt.log.target)
            return set_error(TUPAN_ERR_ARG, "the graph-safe Sakura form takes no guess network or log");
        if (!status) return set_error(TUPAN_ERR_ARG, "the graph-safe Sakura form needs a status word");
        two_pass = true;
    }
    // the far-pair pass has its own geometry (kFarWpt i per thread)
    Launch L = plan_launch(a.ni, a.nj, BLOCK, two_pass ? kFarWpt : 1, TILE, kSakuraTarget * dev.target_wg, 64,
                           TILE);
    size_t nent = static_cast<size_t>(L.nsplit) * ni;
    if (two_pass && nent * kNearSlots > static_cast<size_t>(INT32_MAX)) {
        if (opt.path == 2 || opt.net || opt.log.target || graph)
            return set_error(TUPAN_ERR_ARG, "too many particles for the two-pass Sakura form");
        two_pass = false;
        L = plan_launch(a.ni, a.nj, BLOCK, 1, TILE, kSakuraTarget * dev.target_wg, 64, TILE);
    }
    a.nsplit = L.nsplit;
    a.jchunk = L.jchunk;
    opt.near_pairs = -1;
    const unsigned grid = static_cast<unsigned>(L.nblock_i) * static_cast<unsigned>(L.nsplit);
    const unsigned rg = static_cast<unsigned>((a.ni + 255) / 256);
    if (!two_pass) {
        if (L.nsplit > 1) {
            a.partial = static_cast<double*>(
                workspace(static_cast<size_t>(L.nsplit) * 6 * ni * sizeof(double), s));
            if (!a.partial) return set_error(TUPAN_ERR_ALLOC, "workspace allocation failed");
        }
        timing_begin(s);
        hipLaunchKernelGGL((sakura_kernel_dev<BLOCK, TILE>), dim3(grid), dim3(BLOCK), 0, s, a);
        timing_end(s);
        if (L.nsplit > 1)
            hipLaunchKernelGGL((reduce_kernel<OpSakura>), dim3(rg), dim3(256), 0, s, a);
        return check_launch("sakura");
    }
    const unsigned fgrid = grid;

    // workspace slot 1: partials [nsplit][6][ni], counts, offsets, the
    // overflow list, block sums, totals, the near pass's counter bank and
    // the near j slots
    auto up = [](size_t b) { return (b + 255) & ~static_cast<size_t>(255); };
    const int n = static_cast<int>(nent);
    const int nb = (n + kScanPer - 1) / kScanPer;
    const size_t b_part = up(nent * 6 * sizeof(double));
    const size_t b_cnt = up(nent * sizeof(int));
    const size_t b_blk = up(static_cast<size_t>(nb) * sizeof(int));
    const size_t b_tot = 256;
    const size_t b_bank = up(kBankSlots * kBankStride * sizeof(unsigned long long));
    const size_t b_slots = up(nent * kNearSlots * sizeof(int));
    char* w1 = static_cast<char*>(
        workspace(b_part + 3 * b_cnt + 2 * b_blk + b_tot + b_bank + b_slots, s, 1));
    if (!w1) return set_error(TUPAN_ERR_ALLOC, "workspace allocation failed");
    char* q = w1;
    a.partial = reinterpret_cast<double*>(q);
    q += b_part;
    a.near_count = reinterpret_cast<int*>(q);
    q += b_cnt;
    int* near_off = reinterpret_cast<int*>(q);
    q += b_cnt;
    int* over_list = reinterpret_cast<int*>(q);
    q += b_cnt;
    int* bsum = reinterpret_cast<int*>(q);
    q += b_blk;
    int* bover = reinterpret_cast<int*>(q);
    q += b_blk;
    int* total_dev = reinterpret_cast<int*>(q);
    q += b_tot;
    unsigned long long* bank = (a.counters || opt.fallbacks)
                                   ? reinterpret_cast<unsigned long long*>(q)
                                   : nullptr;
    q += b_bank;
    a.near_j = reinterpret_cast<int*>(q);

    timing_begin(s);
    // the flag as a template argument: no flag tests in the pair loop
    switch (a.flag) {
    case -1:
        hipLaunchKernelGGL((sakura_far_kernel<BLOCK, TILE, kFarWpt, -1>), dim3(fgrid),
                           dim3(BLOCK), 0, s, a);
        break;
    case 1:
        hipLaunchKernelGGL((sakura_far_kernel<BLOCK, TILE, kFarWpt, 1>), dim3(fgrid),
                           dim3(BLOCK), 0, s, a);
        break;
    case -2:
        hipLaunchKernelGGL((sakura_far_kernel<BLOCK, TILE, kFarWpt, -2>), dim3(fgrid),
                           dim3(BLOCK), 0, s, a);
        break;
    case 2:
        hipLaunchKernelGGL((sakura_far_kernel<BLOCK, TILE, kFarWpt, 2>), dim3(fgrid),
                           dim3(BLOCK), 0, s, a);
        break;
    default:
        hipLaunchKernelGGL((sakura_far_kernel<BLOCK, TILE, kFarWpt, 0>), dim3(fgrid),
                           dim3(BLOCK), 0, s, a);
        break;
    }
    // timed as two intervals: the far-pair kernel, then the rest of the call
    timing_end(s);
    timing_begin(s);
    hipLaunchKernelGGL(near_block_sums_kernel, dim3(nb), dim3(kScanThreads), 0, s, n,
                       a.near_count, bsum, bover);
    hipLaunchKernelGGL(near_block_offsets_kernel, dim3(1), dim3(kScanThreads), 0, s, nb, bsum,
                       bover, total_dev, bank);
    hipLaunchKernelGGL(near_offsets_kernel, dim3(nb), dim3(kScanThreads), 0, s, n,
                       a.near_count, bsum, near_off, total_dev, over_list);
    const unsigned eg = static_cast<unsigned>((n + 255) / 256);
    if (graph) {
        // No host synchronisation: the list gets the caller's capacity, the
        // near passes read the listed counts on the device, and a longer
        // list sets status bit 2 (the call's outputs are then invalid).
        const int cap = opt.graph_capacity > 0 ? opt.graph_capacity : 1;
        const size_t nt = static_cast<size_t>(cap);
        const size_t b_items = up(nt * sizeof(int2));
        const size_t b_con = up(nt * 6 * sizeof(double));
        char* w2 = static_cast<char*>(workspace(b_items + b_con, s, 2));
        if (!w2) return set_error(TUPAN_ERR_ALLOC, "workspace allocation failed");
        int2* items = reinterpret_cast<int2*>(w2);
        double* contrib = reinterpret_cast<double*>(w2 + b_items);
        hipLaunchKernelGGL(near_capacity_kernel, dim3(1), dim3(64), 0, s, total_dev, cap, status);
        hipLaunchKernelGGL(near_compact_kernel, dim3(eg), dim3(256), 0, s, n, a.ni, a.near_count,
                           near_off, a.near_j, items, cap);
        hipLaunchKernelGGL(sakura_near_kernel, dim3((cap + 127) / 128), dim3(128), 0, s, a, cap,
                           total_dev, items, contrib, nullptr, opt.log, bank);
        if (bank)
            hipLaunchKernelGGL(near_bank_flush_kernel, dim3(1), dim3(64), 0, s, bank, a.counters,
                               opt.fallbacks);
        hipLaunchKernelGGL(near_add_kernel, dim3(eg), dim3(256), 0, s, n, a.ni, cap,
                           a.near_count, near_off, contrib, a.partial);
        hipLaunchKernelGGL(near_overflow_kernel, dim3(kOverflowWaves), dim3(kWave), 0, s, a,
                           over_list, 0, total_dev + 1);
        hipLaunchKernelGGL((reduce_kernel<OpSakura>), dim3(rg), dim3(256), 0, s, a);
        timing_end(s);
        opt.near_pairs = -1;
        return check_launch("sakura");
    }
    int counts[2] = {0, 0};
    int rc = check_hip(hipMemcpyAsync(counts, total_dev, 2 * sizeof(int), hipMemcpyDeviceToHost,
                                      s),
                       "near-pair count");
    if (rc == TUPAN_OK) rc = check_hip(hipStreamSynchronize(s), "near-pair count");
    if (rc != TUPAN_OK) return rc;
    const int total = counts[0];
    opt.near_pairs = total;
    if (total > 0) {
        // workspace slot 2: items and contributions
        const size_t nt = static_cast<size_t>(total);
        const size_t b_items = up(nt * sizeof(int2));
        const size_t b_con = up(nt * 6 * sizeof(double));
        char* w2 = static_cast<char*>(workspace(b_items + b_con, s, 2));
        if (!w2) return set_error(TUPAN_ERR_ALLOC, "workspace allocation failed");
        int2* items = reinterpret_cast<int2*>(w2);
        double* contrib = reinterpret_cast<double*>(w2 + b_items);
        hipLaunchKernelGGL(near_compact_kernel, dim3(eg), dim3(256), 0, s, n, a.ni,
                           a.near_count, near_off, a.near_j, items, total);
        const unsigned ig = static_cast<unsigned>((total + 127) / 128);
        if (opt.net) {
            hipLaunchKernelGGL(sakura_near_learned_kernel, dim3(ig), dim3(128), 0, s, a, total,
                               items, contrib, static_cast<const NetPacked*>(opt.net), opt.log,
                               bank);
        } else {
            hipLaunchKernelGGL(sakura_near_kernel, dim3(ig), dim3(128), 0, s, a, total, nullptr,
                               items, contrib, nullptr, opt.log, bank);
        }
        if (bank)
            hipLaunchKernelGGL(near_bank_flush_kernel, dim3(1), dim3(64), 0, s, bank, a.counters,
                               opt.fallbacks);
        hipLaunchKernelGGL(near_add_kernel, dim3(eg), dim3(256), 0, s, n, a.ni, total,
                           a.near_count, near_off, contrib, a.partial);
    }
    if (counts[1] > 0) {
        hipLaunchKernelGGL(near_overflow_kernel, dim3(counts[1]), dim3(kWave), 0, s, a,
                           over_list, counts[1], nullptr);
    }
    hipLaunchKernelGGL((reduce_kernel<OpSakura>), dim3(rg), dim3(256), 0, s, a);
    timing_end(s);
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
    kep::Counters cnt;
    kep::advance_pair(i, j, m, rx, ry, rz, e2, vx, vy, vz, dt, cnt);
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
