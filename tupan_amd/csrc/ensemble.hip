// Ensembles of independent few-body systems in one launch: the Sakura pair
// drift (sakura_kernel.c:5-63) and the potential Phi (phi_kernel.c:5-34)
// evaluated inside each system only.
//
// Layout: the particles of all systems are concatenated in one SoA set;
// offsets[s] .. offsets[s+1] is system s.  A plan packs systems into waves:
// systems are sorted by size and each size bucket gets L lanes per system
// (L in {8, 16, 32, 64}, at least the system size), so a wave holds 64/L
// systems of similar size.  Lane l of a wave serves particle l % L of
// system l / L and loops over that system's j particles.  The lanes of a
// wave step through their j loops together, so each step lasts as long as
// the slowest Kepler solve among the wave's pairs at that step; grouping
// systems of similar size keeps the lanes' loop lengths close.  Every lane
// stays in the wave to the end (inactive lanes compute nothing), so the
// wave reductions of the work counters see all 64 lanes.
//
// With a Kepler-guess network set on the plan (ensemble_plan_set_guess)
// the launch takes the learned form: each j step first evaluates the
// network on the wave's 64 pairs (mlp_wave.h, the same inference as the
// Sakura near pass) and every pair's first solve starts from its guess.
// The network needs all 64 lanes, so each lane runs the wave's longest j
// loop, idle past its own range.
//
// The Sakura drift runs each plan wave as one workgroup of kEnsParts waves:
// wave p takes the p-th of kEnsParts consecutive j ranges of the same lanes,
// and the partial sums are added in part order (((d0 + d1) + d2) + d3), so a
// launch has kEnsParts times as many, shorter waves to hide the solver's
// latency with (4 570 -> 12 416 waves for 4096 systems of U{3..64}).
//
// Compiled with -ffp-contract=off (reference operation order).
#include <algorithm>
#include <cstdint>
#include <vector>

#include "kepler.h"
#include "mlp_wave.h"
#include "tupan_internal.h"

namespace tupan {

struct EnsWave {
    int first;  // index into order[]
    int count;  // systems in this wave
    int lanes;  // lanes per system (L)
    int pad;
};

struct EnsPlanImpl {
    int nsys = 0;
    int ntot = 0;
    int nwaves = 0;
    int min_lanes = 8;
    int* d_offsets = nullptr;
    int* d_order = nullptr;
    EnsWave* d_waves = nullptr;
    int device = -1;
    const NetPacked* net = nullptr;           // Kepler-guess network or null
    unsigned long long* fallbacks = nullptr;  // guess fallback counter or null
};

struct EnsArgs {
    const int* offsets;
    const int* order;
    const EnsWave* waves;
    int nwaves;
    const double* p[8];  // m rx ry rz e2 vx vy vz
    double* out[6];
    double dt;
    long flag;
    unsigned long long* counters;
    int* status;
    const NetPacked* net;
    unsigned long long* fallbacks;
};

__device__ __forceinline__ unsigned int ens_wave_sum(unsigned int v)
{
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
    return v;
}

// Lane -> (particle index, system range); active false for idle lanes.
struct EnsLane {
    bool active;
    int i, j0, j1;
};

__device__ __forceinline__ EnsLane ens_lane(const EnsArgs& a, int wave, int lane)
{
    EnsLane r{false, 0, 0, 0};
    if (wave >= a.nwaves) return r;
    const EnsWave w = a.waves[wave];
    const int slot = lane / w.lanes;
    const int il = lane % w.lanes;
    if (slot >= w.count) return r;
    const int s = a.order[w.first + slot];
    const int b = a.offsets[s];
    const int n = a.offsets[s + 1] - b;
    if (il >= n) return r;
    r.active = true;
    r.i = b + il;
    r.j0 = b;
    r.j1 = b + n;
    return r;
}

constexpr int kEnsParts = 4;

// The i particle of a lane.
struct EnsI {
    double m, x, y, z, e2, vx, vy, vz;
};

// One pair drift of particle i against j (sakura_kernel_common.h:194-243),
// its increments weighted by m_j / (m_i + m_j) added to d.
__device__ __forceinline__ void ens_pair_drift(const EnsArgs& a, const EnsI& I, int j,
                                               const kep::Guess* guess, kep::Counters& cnt,
                                               double d[6])
{
    const double jm = a.p[0][j];
    const double p0[3] = {I.x - a.p[1][j], I.y - a.p[2][j], I.z - a.p[3][j]};
    const double e2 = I.e2 + a.p[4][j];
    const double v0[3] = {I.vx - a.p[5][j], I.vy - a.p[6][j], I.vz - a.p[7][j]};
    const double m = I.m + jm;
    double p[3] = {p0[0], p0[1], p0[2]};
    double v[3] = {v0[0], v0[1], v0[2]};
    kep::drift_twobody(a.dt, a.flag, m, e2, p, v, cnt, guess);
    const double mu = jm / m;
    d[0] += mu * (p[0] - p0[0]);
    d[1] += mu * (p[1] - p0[1]);
    d[2] += mu * (p[2] - p0[2]);
    d[3] += mu * (v[0] - v0[0]);
    d[4] += mu * (v[1] - v0[1]);
    d[5] += mu * (v[2] - v0[2]);
}

// The network inputs of pair (i, j): the first twobody stage's features
// (kepler.h: SolveFeatures, NaN when that stage has no solve setup),
// prepared as mlp_wave.h expects.
__device__ __forceinline__ void ens_guess_inputs(const EnsArgs& a, const EnsI& I, int j,
                                                 const NetPacked* net, float xs[4])
{
    double q[3] = {I.x - a.p[1][j], I.y - a.p[2][j], I.z - a.p[3][j]};
    const double e2 = I.e2 + a.p[4][j];
    const double v0[3] = {I.vx - a.p[5][j], I.vy - a.p[6][j], I.vz - a.p[7][j]};
    const double m = I.m + a.p[0][j];
    const double dts = kep::first_stage(a.dt, a.flag, q, v0);
    double k[5], s0;
    kep::SolveFeatures f;
    if (!kep::solve_setup(dts, m, e2, q, v0, k, s0, &f))
        f.tau = f.sigma = f.kappa = f.eps = __builtin_nanf("");
    prepare_features(f.tau, f.sigma, f.kappa, f.eps, net->mean, net->inv_std, xs);
}

template <bool kLearned>
__device__ __forceinline__ void ensemble_sakura_body(const EnsArgs& a)
{
    __shared__ double part_sum[kEnsParts - 1][6][kWave];
    __shared__ unsigned int part_cnt[kEnsParts][4];
    __shared__ int part_fail[kEnsParts];
    const int wave = static_cast<int>(blockIdx.x);
    const int part = static_cast<int>(threadIdx.x) / kWave;
    const int lane = static_cast<int>(threadIdx.x) % kWave;
    const EnsLane L = ens_lane(a, wave, lane);
    kep::Counters cnt;
    double d[6] = {0, 0, 0, 0, 0, 0};
    int ja = 0, jb = 0;
    EnsI I{};
    if (L.active) {
        const int i = L.i;
        const int nj = L.j1 - L.j0;
        ja = L.j0 + (nj * part) / kEnsParts;
        jb = L.j0 + (nj * (part + 1)) / kEnsParts;
        I = EnsI{a.p[0][i], a.p[1][i], a.p[2][i], a.p[3][i],
                 a.p[4][i], a.p[5][i], a.p[6][i], a.p[7][i]};
    }
    if constexpr (kLearned) {
        __shared__ __align__(16) NetPacked lnet;
        mlp_stage(a.net, &lnet);
        int steps = jb - ja;
#pragma unroll
        for (int off = kWave / 2; off > 0; off >>= 1)
            steps = max(steps, __shfl_xor(steps, off, kWave));
        for (int k = 0; k < steps; ++k) {
            // the weights are read from LDS as the network uses them, not
            // hoisted out of the j loop and kept live across the solves
            __asm__ volatile("" ::: "memory");
            const bool has = ja + k < jb;
            float xs[4] = {0.f, 0.f, 0.f, 0.f};
            if (has) ens_guess_inputs(a, I, ja + k, &lnet, xs);
            // lanes l and l + 32 supply the B operands of items l mod 32
            // and 32 + l mod 32
            float x[2][4];
#pragma unroll
            for (int c = 0; c < 2; ++c) {
#pragma unroll
                for (int q = 0; q < 4; ++q)
                    x[c][q] = __shfl(xs[q], 32 * c + (lane & 31), kWave);
            }
            float y[2];
            mlp_wave_eval(&lnet, x, lane, y);
            const float dl = mlp_delta(&lnet, lane < 32 ? y[0] : y[1]);
            if (has) {
                kep::Guess gs;
                gs.use = dl == dl && dl > -0.9f && dl < 8.0f;
                gs.delta = static_cast<double>(dl);
                ens_pair_drift(a, I, ja + k, &gs, cnt, d);
            }
        }
        const unsigned int fb = ens_wave_sum(L.active ? cnt.fallbacks : 0u);
        if (lane == 0 && fb && a.fallbacks)
            atomicAdd(a.fallbacks, static_cast<unsigned long long>(fb));
    } else {
        for (int j = ja; j < jb; ++j) ens_pair_drift(a, I, j, nullptr, cnt, d);
    }
    if (part > 0) {
#pragma unroll
        for (int k = 0; k < 6; ++k) part_sum[part - 1][k][lane] = d[k];
    }
    const unsigned int s = ens_wave_sum(L.active ? cnt.solves : 0u);
    const unsigned int it = ens_wave_sum(L.active ? cnt.iters : 0u);
    const unsigned int rt = ens_wave_sum(L.active ? cnt.retries : 0u);
    const unsigned int lf = ens_wave_sum(L.active ? cnt.leapfrog : 0u);
    const unsigned int fl = ens_wave_sum(L.active && cnt.fail ? 1u : 0u);
    if (lane == 0) {
        part_cnt[part][0] = s;
        part_cnt[part][1] = it;
        part_cnt[part][2] = rt;
        part_cnt[part][3] = lf;
        part_fail[part] = fl != 0;
    }
    __syncthreads();
    if (part != 0) return;
    if (L.active) {
#pragma unroll
        for (int k = 0; k < 6; ++k) {
            double t = d[k];
#pragma unroll
            for (int q = 0; q < kEnsParts - 1; ++q) t += part_sum[q][k][lane];
            a.out[k][L.i] = t;
        }
    }
    if (lane == 0) {
        unsigned long long c[4] = {0, 0, 0, 0};
        int fail = 0;
#pragma unroll
        for (int q = 0; q < kEnsParts; ++q) {
#pragma unroll
            for (int k = 0; k < 4; ++k) c[k] += part_cnt[q][k];
            fail |= part_fail[q];
        }
        if (a.counters) {
#pragma unroll
            for (int k = 0; k < 4; ++k)
                if (c[k]) atomicAdd(a.counters + k, c[k]);
        }
        if (a.status && fail) atomicOr(a.status, 1);
    }
}

__global__ __launch_bounds__(kWave * kEnsParts) void ensemble_sakura_kernel(EnsArgs a)
{
    ensemble_sakura_body<false>(a);
}

// The learned form held to two waves per SIMD (the network's registers
// would otherwise take it to one).
__global__ __launch_bounds__(kWave * kEnsParts, 2) void ensemble_sakura_learned_kernel(EnsArgs a)
{
    ensemble_sakura_body<true>(a);
}

// Potential of each particle due to its own system (phi_kernel_common.h:7-32),
// FP64 with IEEE division and square root in the reference's order.
__global__ __launch_bounds__(256) void ensemble_phi_kernel(EnsArgs a)
{
    const int gtid = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    const EnsLane L = ens_lane(a, gtid / kWave, gtid % kWave);
    if (!L.active) return;
    const int i = L.i;
    const double ix = a.p[1][i], iy = a.p[2][i], iz = a.p[3][i], ie2 = a.p[4][i];
    double phi = 0;
    for (int j = L.j0; j < L.j1; ++j) {
        const double dx = ix - a.p[1][j];
        const double dy = iy - a.p[2][j];
        const double dz = iz - a.p[3][j];
        const double e2 = ie2 + a.p[4][j];
        const double r2 = dx * dx + dy * dy + dz * dz;
        double q = 1.0 / (r2 + e2);
        q = r2 > 0 ? q : 0.0;
        phi -= a.p[0][j] * __builtin_sqrt(q);
    }
    a.out[0][i] = phi;
}

// ---------------------------------------------------------------------------
// Plans.
// ---------------------------------------------------------------------------
static int lanes_for(int n, int min_lanes)
{
    int L = min_lanes;
    while (L < n && L < 64) L *= 2;
    return L;
}

int ensemble_plan_create(int nsys, const int* h_offsets, int min_lanes, void** plan_out)
{
    if (!plan_out) return set_error(TUPAN_ERR_ARG, "null plan pointer");
    *plan_out = nullptr;
    if (nsys < 0 || (nsys > 0 && !h_offsets)) return set_error(TUPAN_ERR_ARG, "bad offsets");
    if (min_lanes != 8 && min_lanes != 16 && min_lanes != 32 && min_lanes != 64)
        return set_error(TUPAN_ERR_ARG, "min_lanes must be 8, 16, 32 or 64");
    std::vector<int> order(nsys);
    for (int s = 0; s < nsys; ++s) {
        const int n = h_offsets[s + 1] - h_offsets[s];
        if (n < 0 || n > 64) return set_error(TUPAN_ERR_ARG, "system sizes must be in [0, 64]");
        if (h_offsets[s] < 0) return set_error(TUPAN_ERR_ARG, "negative offset");
        order[s] = s;
    }
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
        return (h_offsets[x + 1] - h_offsets[x]) < (h_offsets[y + 1] - h_offsets[y]);
    });
    std::vector<EnsWave> waves;
    int k = 0;
    while (k < nsys) {
        const int n = h_offsets[order[k] + 1] - h_offsets[order[k]];
        const int L = lanes_for(n, min_lanes);
        const int per = kWave / L;
        EnsWave w{k, 0, L, 0};
        while (k < nsys && w.count < per &&
               lanes_for(h_offsets[order[k] + 1] - h_offsets[order[k]], min_lanes) == L) {
            ++w.count;
            ++k;
        }
        waves.push_back(w);
    }
    EnsPlanImpl* p = new EnsPlanImpl();
    p->nsys = nsys;
    p->ntot = nsys > 0 ? h_offsets[nsys] : 0;
    p->nwaves = static_cast<int>(waves.size());
    p->min_lanes = min_lanes;
    (void)hipGetDevice(&p->device);
    int rc = TUPAN_OK;
    if (nsys > 0) {
        rc = check_hip(hipMalloc(&p->d_offsets, sizeof(int) * (nsys + 1)), "plan alloc");
        if (rc == TUPAN_OK) rc = check_hip(hipMalloc(&p->d_order, sizeof(int) * nsys), "plan alloc");
        if (rc == TUPAN_OK)
            rc = check_hip(hipMalloc(&p->d_waves, sizeof(EnsWave) * waves.size()), "plan alloc");
        if (rc == TUPAN_OK)
            rc = check_hip(hipMemcpy(p->d_offsets, h_offsets, sizeof(int) * (nsys + 1),
                                     hipMemcpyHostToDevice), "plan copy");
        if (rc == TUPAN_OK)
            rc = check_hip(hipMemcpy(p->d_order, order.data(), sizeof(int) * nsys,
                                     hipMemcpyHostToDevice), "plan copy");
        if (rc == TUPAN_OK)
            rc = check_hip(hipMemcpy(p->d_waves, waves.data(), sizeof(EnsWave) * waves.size(),
                                     hipMemcpyHostToDevice), "plan copy");
    }
    if (rc != TUPAN_OK) {
        ensemble_plan_destroy(p);
        return rc;
    }
    *plan_out = p;
    return TUPAN_OK;
}

int ensemble_plan_destroy(void* plan)
{
    EnsPlanImpl* p = static_cast<EnsPlanImpl*>(plan);
    if (!p) return TUPAN_OK;
    (void)hipFree(p->d_offsets);
    (void)hipFree(p->d_order);
    (void)hipFree(p->d_waves);
    delete p;
    return TUPAN_OK;
}

int ensemble_plan_info(const void* plan, int* nsys, int* ntot, int* nwaves)
{
    const EnsPlanImpl* p = static_cast<const EnsPlanImpl*>(plan);
    if (!p) return set_error(TUPAN_ERR_ARG, "null plan");
    if (nsys) *nsys = p->nsys;
    if (ntot) *ntot = p->ntot;
    if (nwaves) *nwaves = p->nwaves;
    return TUPAN_OK;
}

int ensemble_plan_set_guess(void* plan, const void* net, unsigned long long* fallbacks)
{
    EnsPlanImpl* p = static_cast<EnsPlanImpl*>(plan);
    if (!p) return set_error(TUPAN_ERR_ARG, "null plan");
    if (reinterpret_cast<uintptr_t>(net) % 16 != 0)
        return set_error(TUPAN_ERR_ARG, "network block must be 16-byte aligned");
    p->net = static_cast<const NetPacked*>(net);
    p->fallbacks = net ? fallbacks : nullptr;
    return TUPAN_OK;
}

static int ens_args(const void* plan, const double* const* in8, EnsArgs& a)
{
    const EnsPlanImpl* p = static_cast<const EnsPlanImpl*>(plan);
    if (!p) return set_error(TUPAN_ERR_ARG, "null plan");
    a = EnsArgs{};
    a.offsets = p->d_offsets;
    a.order = p->d_order;
    a.waves = p->d_waves;
    a.nwaves = p->nwaves;
    a.net = p->net;
    a.fallbacks = p->fallbacks;
    for (int f = 0; f < 8; ++f) {
        if (p->ntot > 0 && !in8[f]) return set_error(TUPAN_ERR_ARG, "null input array");
        a.p[f] = in8[f];
    }
    return TUPAN_OK;
}

int ensemble_sakura_call(const void* plan, const double* const* in8, double dt, long flag,
                         double* const* out6, unsigned long long* counters, int* status,
                         void* stream)
{
    EnsArgs a;
    int rc = ens_args(plan, in8, a);
    if (rc != TUPAN_OK) return rc;
    for (int k = 0; k < 6; ++k) {
        if (!out6[k] && static_cast<const EnsPlanImpl*>(plan)->ntot > 0)
            return set_error(TUPAN_ERR_ARG, "null output array");
        a.out[k] = out6[k];
    }
    a.dt = dt;
    a.flag = flag;
    a.counters = counters;
    a.status = status;
    if (a.nwaves == 0) return TUPAN_OK;
    hipStream_t s = static_cast<hipStream_t>(stream);
    timing_begin(s);
    if (a.net)
        hipLaunchKernelGGL(ensemble_sakura_learned_kernel, dim3(a.nwaves), dim3(kWave * kEnsParts),
                           0, s, a);
    else
        hipLaunchKernelGGL(ensemble_sakura_kernel, dim3(a.nwaves), dim3(kWave * kEnsParts),
                           0, s, a);
    timing_end(s);
    return check_launch("ensemble_sakura");
}

int ensemble_phi_call(const void* plan, const double* const* in8, double* phi, void* stream)
{
    EnsArgs a;
    int rc = ens_args(plan, in8, a);
    if (rc != TUPAN_OK) return rc;
    if (!phi && static_cast<const EnsPlanImpl*>(plan)->ntot > 0)
        return set_error(TUPAN_ERR_ARG, "null output array");
    a.out[0] = phi;
    if (a.nwaves == 0) return TUPAN_OK;
    const unsigned blocks = static_cast<unsigned>((a.nwaves + 3) / 4);
    hipLaunchKernelGGL(ensemble_phi_kernel, dim3(blocks), dim3(256), 0,
                       static_cast<hipStream_t>(stream), a);
    return check_launch("ensemble_phi");
}

}  // namespace tupan
