// All-pairs gravity kernels for gfx950: launch plumbing, tile shapes, the
// tuning table and the dispatch of the pair interactions of ops.h.
//
// This file is compiled with -ffp-contract=off: every fused multiply-add in
// ops.h is explicit.
#include <mutex>
#include <vector>

#include "ops.h"
#include "pn.h"
#include "tupan_internal.h"

namespace tupan {

// ---------------------------------------------------------------------------
// Launch plumbing and tile configurations.
//
// Each interaction is compiled for a few tile shapes (BLOCK threads, WPT
// i-particles per thread, TILE j-slots per LDS tile, UNROLL j per unrolled
// step).  A run-time table, filled from the autotuner's cache, picks the
// shape and the workgroups-per-CU target per kernel, mode and size bucket.
// ---------------------------------------------------------------------------

// Device memory a caller lends for the j-split partial sums instead of the
// library's growable workspace: `doubles` values at `p`.  The launch plan
// then caps its j-split count so the partials fit, and the call records no
// timing events.
struct FixedScratch {
    double* p;
    size_t doubles;
};

// The FP64 pass of the ops whose FP32 overflow fallback runs after the call.
template <class Op>
static void launch_post_fallback(const typename Op::Args& a, hipStream_t stream)
{
    if constexpr (HasPostFallback<Op>::value) {
        if (a.nj > 0) {
            const unsigned g = static_cast<unsigned>((a.ni + 255) / 256);
            hipLaunchKernelGGL((fallback_kernel<Op>), dim3(g), dim3(256), 0, stream, a);
        }
    }
}

// ---------------------------------------------------------------------------
// Coincidence scan (split AccJerk): does an i share its position with a j
// of different velocity?  Position means all three FP64 components equal,
// the reference's r^2 = 0 (smoothing.h: such a pair is dropped), and such a
// pair, unlike a particle with itself, would add a jerk term if its tile ran
// unmasked (ops.h: OpSplitAccJerkT::unmask_allowed).  The j positions go
// into a hash table (open addressing, linear probing, a power of two of at
// least four times nj entries), every i probes it, and the flag stays 0 when no
// such pair exists and every j's zero-separation term is finite
// (Op::self_safe_j; bit 1 otherwise).  A NaN position equals nothing; its
// pairs give non-finite sums and take the FP64 fallback either way.  Calls
// below kCoincideMinPairs pairs skip the scan and keep the mask.
// ---------------------------------------------------------------------------
constexpr double kCoincideMinPairs = 268435456.0;  // 2^28

template <class Op, class = void>
struct HasCoincideScan : std::false_type {};
template <class Op>
struct HasCoincideScan<Op, std::void_t<decltype(Op::kCoincideScan)>>
    : std::integral_constant<bool, Op::kCoincideScan> {};

__device__ __forceinline__ unsigned long long coincide_bits(double v)
{
    return static_cast<unsigned long long>(__double_as_longlong(v == 0.0 ? 0.0 : v));
}

__device__ __forceinline__ unsigned coincide_hash(double x, double y, double z, unsigned mask)
{
    unsigned long long h = coincide_bits(x) * 0x9E3779B97F4A7C15ull;
    h ^= (coincide_bits(y) + 0x632BE59BD9B4E019ull) * 0xC2B2AE3D27D4EB4Full;
    h ^= (coincide_bits(z) + 0x165667B19E3779F9ull) * 0xD6E8FEB86659FD93ull;
    h ^= h >> 31;
    h *= 0x94D049BB133111EBull;
    h ^= h >> 29;
    return static_cast<unsigned>(h) & mask;
}

__global__ __launch_bounds__(256) void coincide_clear_kernel(int* table, unsigned size, int* flag)
{
    const unsigned k = blockIdx.x * 256u + threadIdx.x;
    if (k < size) table[k] = -1;
    if (k == 0) *flag = 0;
}

// j and k (both j-particles) at one position with different velocities.
__device__ __forceinline__ bool coincide_pair(const PairArgs<double>& a, int j, int k)
{
    return a.jp[1][k] == a.jp[1][j] && a.jp[2][k] == a.jp[2][j] && a.jp[3][k] == a.jp[3][j] &&
           !(a.jp[5][k] == a.jp[5][j] && a.jp[6][k] == a.jp[6][j] && a.jp[7][k] == a.jp[7][j]);
}

// With `same` (the i-set is the j-set) the insertion is the whole scan:
// of two particles at one position, the one inserted second passes the
// first one's entry on its way to a free one, and compares with it.
template <class Op>
__global__ __launch_bounds__(256) void coincide_insert_kernel(PairArgs<double> a, int* table,
                                                              unsigned mask, int* flag, int same)
{
    const int j = static_cast<int>(blockIdx.x * 256u + threadIdx.x);
    if (j >= a.nj) return;
    if (!Op::self_safe_j(a, j)) atomicOr(flag, 2);
    unsigned h = coincide_hash(a.jp[1][j], a.jp[2][j], a.jp[3][j], mask);
    for (unsigned probe = 0; probe <= mask; ++probe) {
        const int k = atomicCAS(&table[h], -1, j);
        if (k == -1) return;
        if (same && coincide_pair(a, j, k)) {
            atomicOr(flag, 1);
            return;
        }
        h = (h + 1) & mask;
    }
    atomicOr(flag, 1);  // no free entry (not reached at quarter load): keep the mask
}

__global__ __launch_bounds__(256) void coincide_probe_kernel(PairArgs<double> a,
                                                             const int* table, unsigned mask,
                                                             int* flag)
{
    const int i = static_cast<int>(blockIdx.x * 256u + threadIdx.x);
    if (i >= a.ni) return;
    const double x = a.ip[1][i], y = a.ip[2][i], z = a.ip[3][i];
    unsigned h = coincide_hash(x, y, z, mask);
    for (unsigned probe = 0; probe <= mask; ++probe) {
        const int j = table[h];
        if (j < 0) return;
        if (a.jp[1][j] == x && a.jp[2][j] == y && a.jp[3][j] == z &&
            !(a.jp[5][j] == a.ip[5][i] && a.jp[6][j] == a.ip[6][i] && a.jp[7][j] == a.ip[7][i])) {
            atomicOr(flag, 1);
            return;
        }
        h = (h + 1) & mask;
    }
}

// Entries of the coincidence table for nj j-particles: a quarter full at
// most, so an insertion or a probe ends within ~1.4 entries on average.
static unsigned coincide_entries(int nj)
{
    unsigned size = 256;
    while (size < 4u * static_cast<unsigned>(nj)) size <<= 1;
    return size;
}

template <class Op, int BLOCK, int WPT, int TILE, int UNROLL>
static int run_pairwise(typename Op::Args a, hipStream_t stream, int wg_per_cu,
                        const FixedScratch* fixed)
{
    if (a.ni <= 0) return 0;
    if (!a.exact && a.nj > kSmallMinJ && a.nj <= kSmallMaxJ &&
        static_cast<long long>(a.ni) * a.nj <= kSmallPairs) {
        a.nsplit = 1;
        a.jchunk = a.nj;
        a.partial = nullptr;
        constexpr int ipb = kSmallBlock / kSmallGroups;
        const unsigned grid = static_cast<unsigned>((a.ni + ipb - 1) / ipb);
        if (!fixed) timing_begin(stream);
        hipLaunchKernelGGL((pair_small_kernel<Op, kSmallBlock, kSmallGroups, small_tile<Op>()>),
                           dim3(grid), dim3(kSmallBlock), 0, stream, a);
        if (!fixed) timing_end(stream);
        launch_post_fallback<Op>(a, stream);
        return check_launch("pairwise");
    }
    const Device& dev = current_device();
    const int target = (wg_per_cu > 0 ? wg_per_cu : 4) * dev.cu_count;
    // j-chunks of whole tiles for a set of one tile, 32-particle chunks for
    // larger sets in the modes that sum in any order, so that a small system
    // (N ~ 1e3) fills the chip with j-splits.  ieee64 never cuts a j-range
    // below 256 particles, so any set of at most 256 is one split (the
    // reference's summation order) whatever shape is tuned.
    const int min_chunk = min_jchunk(a.nj, TILE, a.exact != 0);
    // at most 64 j-splits; with lent scratch, at most as many as it holds
    int max_split = 64;
    if (fixed) {
        const size_t per_split = static_cast<size_t>(Op::kNacc) * static_cast<size_t>(a.ni);
        const size_t fit = fixed->doubles / per_split;
        if (fit < static_cast<size_t>(max_split)) max_split = fit > 1 ? static_cast<int>(fit) : 1;
    }
    Launch L = plan_launch(a.ni, a.nj, BLOCK, WPT, TILE, target, max_split, min_chunk);
    a.nsplit = L.nsplit;
    a.jchunk = L.jchunk;
    a.partial = nullptr;
    a.coincide = nullptr;
    // the coincidence table and flag follow the partials in the workspace
    bool scan = false;
    if constexpr (HasCoincideScan<Op>::value)
        scan = !fixed && static_cast<double>(a.ni) * static_cast<double>(a.nj) >= kCoincideMinPairs;
    const size_t part_bytes =
        L.nsplit > 1 ? static_cast<size_t>(L.nsplit) * Op::kNacc * a.ni * sizeof(double) : 0;
    const size_t part_round = (part_bytes + 255) / 256 * 256;
    const unsigned table_n = scan ? coincide_entries(a.nj) : 0u;
    const size_t scan_bytes = scan ? static_cast<size_t>(table_n) * sizeof(int) + 256 : 0;
    if (L.nsplit > 1 || scan) {
        if (fixed) {
            if (part_bytes > fixed->doubles * sizeof(double))
                return set_error(TUPAN_ERR_ARG, "j-split partials exceed the fixed scratch");
            a.partial = fixed->p;
        } else {
            char* w = static_cast<char*>(workspace(part_round + scan_bytes, stream));
            if (!w) return set_error(TUPAN_ERR_ALLOC, "workspace allocation failed");
            if (L.nsplit > 1) a.partial = reinterpret_cast<double*>(w);
            if constexpr (HasCoincideScan<Op>::value) {
                if (scan) {
                    int* flag = reinterpret_cast<int*>(w + part_round);
                    int* table = flag + 64;
                    const unsigned mask = table_n - 1;
                    hipLaunchKernelGGL(coincide_clear_kernel, dim3((table_n + 255) / 256),
                                       dim3(256), 0, stream, table, table_n, flag);
                    bool same = a.ni == a.nj;
                    for (int f : {1, 2, 3, 5, 6, 7}) same = same && a.ip[f] == a.jp[f];
                    hipLaunchKernelGGL(coincide_insert_kernel<Op>,
                                       dim3(static_cast<unsigned>((a.nj + 255) / 256)), dim3(256),
                                       0, stream, a, table, mask, flag, same ? 1 : 0);
                    if (!same)
                        hipLaunchKernelGGL(coincide_probe_kernel,
                                           dim3(static_cast<unsigned>((a.ni + 255) / 256)),
                                           dim3(256), 0, stream, a,
                                           static_cast<const int*>(table), mask, flag);
                    a.coincide = flag;
                }
            }
        }
    }
    const unsigned grid = static_cast<unsigned>(L.nblock_i) * static_cast<unsigned>(L.nsplit);
    if (!fixed) timing_begin(stream);
    hipLaunchKernelGGL((pair_kernel<Op, BLOCK, WPT, TILE, UNROLL>), dim3(grid), dim3(BLOCK), 0,
                       stream, a);
    if (!fixed) timing_end(stream);
    if (L.nsplit > 1) {
        if (a.ni <= kWideReduceMaxNi) {
            const unsigned rg = static_cast<unsigned>((static_cast<size_t>(a.ni) * 8 + 255) / 256);
            hipLaunchKernelGGL((reduce_wide_kernel<Op, 8>), dim3(rg), dim3(256), 0, stream, a);
        } else {
            const unsigned rg = static_cast<unsigned>((a.ni + 255) / 256);
            hipLaunchKernelGGL((reduce_kernel<Op>), dim3(rg), dim3(256), 0, stream, a);
        }
    }
    launch_post_fallback<Op>(a, stream);
    return check_launch("pairwise");
}

template <class Op>
using Launcher = int (*)(typename Op::Args, hipStream_t, int, const FixedScratch*);

// Shapes for the light interactions (Phi, Acc, AccJerk, Tstep, split); the
// last, one wave per workgroup, is for small systems.
template <class Op, int S>
static int run_light(typename Op::Args a, hipStream_t s, int wg, const FixedScratch* f)
{
    return run_pairwise<Op, kLightShapeDims[S][0], kLightShapeDims[S][1], kLightShapeDims[S][2],
                        kLightShapeDims[S][3]>(a, s, wg, f);
}
#define TUPAN_LIGHT_SHAPES(Op)                                                   \
    {                                                                            \
        &run_light<Op, 0>, &run_light<Op, 1>, &run_light<Op, 2>, &run_light<Op, 3>, \
            &run_light<Op, 4>, &run_light<Op, 5>                                 \
    }
// Shapes for SnapCrackle (twice the registers per i).
#define TUPAN_HEAVY_SHAPES(Op)                                                   \
    {                                                                            \
        &run_pairwise<Op, 256, 1, 128, 2>, &run_pairwise<Op, 128, 1, 128, 2>,    \
            &run_pairwise<Op, 256, 2, 128, 1>, &run_pairwise<Op, 64, 1, 64, 2>,  \
            &run_pairwise<Op, 256, 1, 256, 4>, &run_pairwise<Op, 64, 1, 32, 1>   \
    }
// Shapes for PNAcc (a few hundred operations and ~100 VGPRs per pair).
#define TUPAN_PN_SHAPES(Op)                                                      \
    {                                                                            \
        &run_pairwise<Op, 256, 1, 128, 1>, &run_pairwise<Op, 128, 1, 128, 1>,    \
            &run_pairwise<Op, 256, 1, 256, 2>, &run_pairwise<Op, 64, 1, 64, 1>,  \
            &run_pairwise<Op, 512, 1, 256, 1>, &run_pairwise<Op, 64, 1, 32, 1>   \
    }
constexpr int kShapes = kNumShapes;
static const int (&kLightShape)[kShapes][4] = kLightShapeDims;
static const int kHeavyShape[kShapes][4] = {
    {256, 1, 128, 2}, {128, 1, 128, 2}, {256, 2, 128, 1}, {64, 1, 64, 2}, {256, 1, 256, 4},
    {64, 1, 32, 1}};
static const int kPnShape[kShapes][4] = {
    {256, 1, 128, 1}, {128, 1, 128, 1}, {256, 1, 256, 2}, {64, 1, 64, 1}, {512, 1, 256, 1},
    {64, 1, 32, 1}};

// Size buckets by pair count ni * nj.
constexpr int kBuckets = 3;
static int size_bucket(int ni, int nj)
{
    const double pairs = static_cast<double>(ni) * static_cast<double>(nj);
    if (pairs < 4194304.0) return 0;        // < 2^22
    if (pairs < 1073741824.0) return 1;     // < 2^30
    return 2;
}

// Tuning table: [kernel][mode slot][bucket] -> (shape, workgroups per CU).
// Mode slots: 0 fp64, 1 split32, 2 ieee64, 3 float build.
struct TuneEntry {
    int shape = 0;
    int wg_per_cu = 4;
};
static TuneEntry g_tune[kNumKernels][4][kBuckets];

static int mode_slot(int mode, bool f32)
{
    if (f32) return 3;
    if (mode == TUPAN_MODE_SPLIT32) return 1;
    if (mode == TUPAN_MODE_IEEE64) return 2;
    return 0;
}

template <class Op>
static int launch_tuned(const Launcher<Op> (&table)[kShapes], int kernel, int slot,
                        typename Op::Args& a, hipStream_t s, const FixedScratch* fixed)
{
    const TuneEntry& t = g_tune[kernel][slot][size_bucket(a.ni, a.nj)];
    const int shape = (t.shape >= 0 && t.shape < kShapes) ? t.shape : 0;
    return table[shape](a, s, t.wg_per_cu, fixed);
}

#define TUPAN_RUN(Op, SHAPES, kernel, slot)                                      \
    do {                                                                         \
        static const Launcher<Op> table[kShapes] = SHAPES(Op);                  \
        return launch_tuned<Op>(table, kernel, slot, a, s, fixed);               \
    } while (0)

// Kernel selection by name and mode.
template <class T>
static int dispatch(int kernel, int mode, PairArgs<T>& a, hipStream_t s,
                    const FixedScratch* fixed);

template <>
int dispatch<double>(int kernel, int mode, PairArgs<double>& a, hipStream_t s,
                     const FixedScratch* fixed)
{
    const int slot = mode_slot(mode, false);
    switch (kernel) {
    case K_PHI:
        if (mode == TUPAN_MODE_IEEE64) TUPAN_RUN(OpPhi<Ieee64>, TUPAN_LIGHT_SHAPES, K_PHI, slot);
        TUPAN_RUN(OpPhi<Fast64>, TUPAN_LIGHT_SHAPES, K_PHI, slot);
    case K_ACC:
        if (mode == TUPAN_MODE_SPLIT32) TUPAN_RUN(OpSplitAcc, TUPAN_LIGHT_SHAPES, K_ACC, slot);
        if (mode == TUPAN_MODE_IEEE64) TUPAN_RUN(OpAcc<Ieee64>, TUPAN_LIGHT_SHAPES, K_ACC, slot);
        TUPAN_RUN(OpAcc<Fast64>, TUPAN_LIGHT_SHAPES, K_ACC, slot);
    case K_ACC_JERK:
        if (mode == TUPAN_MODE_SPLIT32)
            TUPAN_RUN(OpSplitAccJerk, TUPAN_LIGHT_SHAPES, K_ACC_JERK, slot);
        if (mode == TUPAN_MODE_IEEE64)
            TUPAN_RUN(OpAccJerk<Ieee64>, TUPAN_LIGHT_SHAPES, K_ACC_JERK, slot);
        TUPAN_RUN(OpAccJerk<Fast64>, TUPAN_LIGHT_SHAPES, K_ACC_JERK, slot);
    case K_SNAP_CRACKLE:
        if (mode == TUPAN_MODE_IEEE64)
            TUPAN_RUN(OpSnapCrackle<Ieee64>, TUPAN_HEAVY_SHAPES, K_SNAP_CRACKLE, slot);
        TUPAN_RUN(OpSnapCrackle<Fast64>, TUPAN_HEAVY_SHAPES, K_SNAP_CRACKLE, slot);
    case K_TSTEP:
        if (mode == TUPAN_MODE_SPLIT32) TUPAN_RUN(OpSplitTstep, TUPAN_LIGHT_SHAPES, K_TSTEP, slot);
        if (mode == TUPAN_MODE_IEEE64) TUPAN_RUN(OpTstep<Ieee64>, TUPAN_LIGHT_SHAPES, K_TSTEP, slot);
        TUPAN_RUN(OpTstep<Fast64>, TUPAN_LIGHT_SHAPES, K_TSTEP, slot);
    case K_NREG_X:
        if (mode == TUPAN_MODE_IEEE64) TUPAN_RUN(OpNregX<Ieee64>, TUPAN_LIGHT_SHAPES, K_NREG_X, slot);
        TUPAN_RUN(OpNregX<Fast64>, TUPAN_LIGHT_SHAPES, K_NREG_X, slot);
    case K_NREG_V:
        if (mode == TUPAN_MODE_IEEE64) TUPAN_RUN(OpNregV<Ieee64>, TUPAN_LIGHT_SHAPES, K_NREG_V, slot);
        TUPAN_RUN(OpNregV<Fast64>, TUPAN_LIGHT_SHAPES, K_NREG_V, slot);
    case K_PNACC:
        if (mode == TUPAN_MODE_SPLIT32) TUPAN_RUN(OpSplitPNAcc, TUPAN_PN_SHAPES, K_PNACC, slot);
        if (mode == TUPAN_MODE_IEEE64) TUPAN_RUN(OpPNAcc<Ieee64>, TUPAN_PN_SHAPES, K_PNACC, slot);
        TUPAN_RUN(OpPNAcc<Fast64>, TUPAN_PN_SHAPES, K_PNACC, slot);
    default:
        return set_error(TUPAN_ERR_ARG, "unknown pairwise kernel");
    }
}

template <>
int dispatch<float>(int kernel, int, PairArgs<float>& a, hipStream_t s,
                    const FixedScratch* fixed)
{
    const int slot = mode_slot(0, true);
    switch (kernel) {
    case K_PHI: TUPAN_RUN(OpPhi<Fast32>, TUPAN_LIGHT_SHAPES, K_PHI, slot);
    case K_ACC: TUPAN_RUN(OpAcc<Fast32>, TUPAN_LIGHT_SHAPES, K_ACC, slot);
    case K_ACC_JERK: TUPAN_RUN(OpAccJerk<Fast32>, TUPAN_LIGHT_SHAPES, K_ACC_JERK, slot);
    case K_SNAP_CRACKLE:
        TUPAN_RUN(OpSnapCrackle<Fast32>, TUPAN_HEAVY_SHAPES, K_SNAP_CRACKLE, slot);
    case K_TSTEP: TUPAN_RUN(OpTstep<Fast32>, TUPAN_LIGHT_SHAPES, K_TSTEP, slot);
    case K_NREG_X: TUPAN_RUN(OpNregX<Fast32>, TUPAN_LIGHT_SHAPES, K_NREG_X, slot);
    case K_NREG_V: TUPAN_RUN(OpNregV<Fast32>, TUPAN_LIGHT_SHAPES, K_NREG_V, slot);
    case K_PNACC: TUPAN_RUN(OpPNAcc<Fast32>, TUPAN_PN_SHAPES, K_PNACC, slot);
    default: return set_error(TUPAN_ERR_ARG, "unknown pairwise kernel");
    }
}

int tune_set(int kernel, int slot, int bucket, int shape, int wg_per_cu)
{
    if (kernel < 0 || kernel >= kNumKernels || slot < 0 || slot > 3 || bucket < 0 ||
        bucket >= kBuckets || shape < 0 || shape >= kShapes || wg_per_cu < 1 || wg_per_cu > 64)
        return set_error(TUPAN_ERR_ARG, "tuning entry out of range");
    g_tune[kernel][slot][bucket].shape = shape;
    g_tune[kernel][slot][bucket].wg_per_cu = wg_per_cu;
    return TUPAN_OK;
}

int tune_get(int kernel, int slot, int bucket, int* shape, int* wg_per_cu)
{
    if (kernel < 0 || kernel >= kNumKernels || slot < 0 || slot > 3 || bucket < 0 ||
        bucket >= kBuckets)
        return set_error(TUPAN_ERR_ARG, "tuning entry out of range");
    if (shape) *shape = g_tune[kernel][slot][bucket].shape;
    if (wg_per_cu) *wg_per_cu = g_tune[kernel][slot][bucket].wg_per_cu;
    return TUPAN_OK;
}

int tune_shape(int kernel, int shape, int* block, int* wpt, int* tile, int* unroll)
{
    if (kernel < 0 || kernel >= kNumKernels || shape < 0 || shape >= kShapes)
        return set_error(TUPAN_ERR_ARG, "shape out of range");
    const int* d = kernel == K_SNAP_CRACKLE ? kHeavyShape[shape]
                   : kernel == K_PNACC      ? kPnShape[shape]
                                            : kLightShape[shape];
    if (block) *block = d[0];
    if (wpt) *wpt = d[1];
    if (tile) *tile = d[2];
    if (unroll) *unroll = d[3];
    return TUPAN_OK;
}

namespace {
int noalloc_scratch(double** out);
}

// `fixed`: null for the growable workspace; else lent scratch, whose device
// address is looked up (noalloc_scratch) once the arguments have passed
// their checks.
template <class T>
static int pairwise_impl(int kernel, int mode, size_t ni, const T* const* ip, size_t nj,
                         const T* const* jp, T* const* out, T eta, void* stream,
                         const PnParams* pn, FixedScratch* fixed)
{
    const int nin = kernel_in_fields(kernel);
    const int nout = kernel_out_fields(kernel);
    if (nin < 0) return set_error(TUPAN_ERR_ARG, "unknown pairwise kernel");
    if (ni > static_cast<size_t>(INT32_MAX) || nj > static_cast<size_t>(INT32_MAX))
        return set_error(TUPAN_ERR_ARG, "particle count exceeds 2^31-1");
    if (ni == 0) return 0;
    PairArgs<T> a{};
    for (int f = 0; f < nin; ++f) {
        if (!ip[f]) return set_error(TUPAN_ERR_ARG, "null i-array");
        a.ip[f] = ip[f];
        if (nj > 0 && !jp[f]) return set_error(TUPAN_ERR_ARG, "null j-array");
        a.jp[f] = jp[f];
    }
    for (int k = 0; k < nout; ++k) {
        if (!out[k]) return set_error(TUPAN_ERR_ARG, "null output array");
        a.out[k] = out[k];
    }
    a.ni = static_cast<int>(ni);
    a.nj = static_cast<int>(nj);
    a.eta = eta;
    a.exact = mode == TUPAN_MODE_IEEE64 ? 1 : 0;
    if (kernel == K_PNACC) {
        if (!pn) return set_error(TUPAN_ERR_ARG, "PNAcc needs the PN order and 1/c^k");
        // pn_terms.h:503-548: order 2 and 3 give 1PN, 4 adds 2PN, 5 2.5PN,
        // 6 3PN, 7 and above 3.5PN; below 2 every pair term is zero, so the
        // call runs with no j-particles and stores zeros.
        a.order = pn->order > 7 ? 7 : static_cast<int>(pn->order);
        a.clight[0] = static_cast<T>(pn->inv[1]);
        a.clight[1] = static_cast<T>(pn->inv[3]);
        a.clight[2] = static_cast<T>(pn->inv[4]);
        a.clight[3] = static_cast<T>(pn->inv[5]);
        a.clight[4] = static_cast<T>(pn->inv[6]);
        if (a.order < 2) a.nj = 0;
    }
    if (mode == TUPAN_MODE_SPLIT32 && kernel != K_ACC && kernel != K_ACC_JERK &&
        kernel != K_TSTEP && kernel != K_PNACC)
        mode = TUPAN_MODE_FP64;  // split FP32 exists for the forces, Tstep and PNAcc
    if (fixed && !fixed->p) {
        const int rc = noalloc_scratch(&fixed->p);
        if (rc != TUPAN_OK) return rc;
    }
    return dispatch<T>(kernel, mode, a, static_cast<hipStream_t>(stream), fixed);
}

template <class T>
int pairwise_call(int kernel, int mode, size_t ni, const T* const* ip, size_t nj,
                  const T* const* jp, T* const* out, T eta, void* stream, const PnParams* pn)
{
    return pairwise_impl<T>(kernel, mode, ni, ip, nj, jp, out, eta, stream, pn, nullptr);
}

// ---------------------------------------------------------------------------
// tupan_pairwise_noalloc: the pairwise kernels with their j-split partial
// sums in a fixed device array instead of the growable workspace, so the
// call allocates nothing.  Each device that loads this library holds its
// own copy of `scratch` (64 MiB) until the process ends.  In host code
// `scratch` names a host-side stand-in, never the device array: the device
// address comes from hipGetSymbolAddress, taken once per device with that
// device current, and only that address reaches kernels.
// ---------------------------------------------------------------------------
constexpr size_t kNoallocDoubles = TUPAN_NOALLOC_SCRATCH_BYTES / sizeof(double);
static __device__ __align__(256) double scratch[kNoallocDoubles];

namespace {
constexpr int kMaxNoallocDevices = 64;
std::mutex g_noalloc_mutex;
double* g_noalloc_ptr[kMaxNoallocDevices] = {};

// Device address of `scratch` on the current device (looked up on the first
// call on that device).
int noalloc_scratch(double** out)
{
    int id = 0;
    int rc = check_hip(hipGetDevice(&id), "hipGetDevice");
    if (rc != TUPAN_OK) return rc;
    if (id < 0 || id >= kMaxNoallocDevices)
        return set_error(TUPAN_ERR_ARG, "device index beyond the no-alloc table (64 devices)");
    std::lock_guard<std::mutex> lock(g_noalloc_mutex);
    if (!g_noalloc_ptr[id]) {
        void* ptr = nullptr;
        rc = check_hip(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
        if (rc != TUPAN_OK) return rc;
        if (!ptr) return set_error(TUPAN_ERR_HIP, "hipGetSymbolAddress(scratch) gave no address");
        g_noalloc_ptr[id] = static_cast<double*>(ptr);
    }
    *out = g_noalloc_ptr[id];
    return TUPAN_OK;
}
}  // namespace

// Mean relative error of v_rsq_f32, raw and after the Newton step the split
// kernels can build with, over every float in [1, 4) (2^24 inputs, both
// exponent parities, weight 1/x).  Each 256-thread block lies in one binade
// and writes its weighted sums; the host adds the blocks in order.
__global__ __launch_bounds__(256) void rsq_bias_kernel(double* part)
{
    __shared__ double sh[3][256];
    const int t = static_cast<int>(threadIdx.x);
    const unsigned k = blockIdx.x * 256u + static_cast<unsigned>(t);
    const float x = __uint_as_float(0x3f800000u + k);
    const double exact = 1.0 / __builtin_sqrt(static_cast<double>(x));
    const double w = 1.0 / static_cast<double>(x);
    const float raw = __builtin_amdgcn_rsqf(x);
    const float ref = rsqrt_refined(x);
    sh[0][t] = w * ((static_cast<double>(raw) - exact) / exact);
    sh[1][t] = w * ((static_cast<double>(ref) - exact) / exact);
    sh[2][t] = w;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (t < h) {
            for (int c = 0; c < 3; ++c) sh[c][t] += sh[c][t + h];
        }
        __syncthreads();
    }
    if (t < 3) part[3 * blockIdx.x + t] = sh[t][0];
}

int rsq_bias_probe(double* out)
{
    constexpr int kBlocks = (1 << 24) / 256;
    double* d = nullptr;
    if (hipMalloc(&d, sizeof(double) * 3 * kBlocks) != hipSuccess) {
        (void)hipGetLastError();
        return set_error(TUPAN_ERR_ALLOC, "rsq probe allocation failed");
    }
    hipLaunchKernelGGL(rsq_bias_kernel, dim3(kBlocks), dim3(256), 0, nullptr, d);
    int rc = check_launch("rsq_bias_kernel");
    std::vector<double> h(3 * kBlocks);
    if (rc == TUPAN_OK)
        rc = check_hip(hipMemcpy(h.data(), d, sizeof(double) * h.size(), hipMemcpyDeviceToHost),
                       "rsq probe copy");
    (void)hipFree(d);
    if (rc != TUPAN_OK) return rc;
    for (int b = 0; b < 2; ++b) {
        double raw = 0, ref = 0, w = 0;
        for (int blk = b * kBlocks / 2; blk < (b + 1) * kBlocks / 2; ++blk) {
            raw += h[3 * blk];
            ref += h[3 * blk + 1];
            w += h[3 * blk + 2];
        }
        out[b] = raw / w;
        out[2 + b] = ref / w;
    }
    out[4] = kRsqBias;
    return 0;
}

int pairwise_noalloc_call(int kernel, int mode, size_t ni, const double* const* ip, size_t nj,
                          const double* const* jp, double* const* out, double eta, void* stream)
{
    FixedScratch fs{nullptr, kNoallocDoubles};
    return pairwise_impl<double>(kernel, mode, ni, ip, nj, jp, out, eta, stream, nullptr, &fs);
}

template int pairwise_call<double>(int, int, size_t, const double* const*, size_t,
                                   const double* const*, double* const*, double, void*,
                                   const PnParams*);
template int pairwise_call<float>(int, int, size_t, const float* const*, size_t,
                                  const float* const*, float* const*, float, void*,
                                  const PnParams*);

}  // namespace tupan
