// All-pairs tile engine for gfx950.
//
// One workgroup owns BLOCK*WPT i-particles (WPT per thread, held in VGPRs)
// and one contiguous range of j-particles (a "j-split").  The j range is
// streamed through LDS in tiles of TILE slots: the workgroup loads a tile
// with coalesced SoA reads from HBM/L2, converts it to the interaction's
// slot format (AoS, 16-byte aligned, so each wave reads a slot with 128-bit
// broadcast LDS reads), and every thread applies the slot to each of its WPT
// i-particles.  With nsplit > 1 every workgroup writes its partial
// accumulators to a scratch array and a second kernel combines the splits in
// a fixed order, so results are deterministic run to run.
//
// An interaction policy `Op` supplies:
//   Args                        kernel argument block
//   Slot, IReg, Acc, Frame      LDS slot, i registers, accumulators, the
//                               per-workgroup frame (centre for split modes)
//   kNacc                       number of FP64 partial accumulators
//   kFlush                      j count between Acc flushes (split modes)
//   frame(a, ibase)             per-workgroup frame
//   load_i(a, fr, i, r)         i registers
//   load_j(a, fr, j, s)         j slot
//   zero(acc), interact(r, s, acc), flush(acc)
//   get(acc, k)                 FP64 value of accumulator k (after flush)
//   combine(tot, part)          merge two partial accumulator sets
//   store(a, i, tot)            final outputs from combined accumulators
#pragma once

#include <type_traits>
#include <utility>

#include "common.h"

namespace tupan {

// Ops that declare kUniformE2 take the pair softening as an argument on
// tiles where it is the same for every pair (see pair_kernel).
template <class Op, class = void>
struct HasUniformE2 : std::false_type {};
template <class Op>
struct HasUniformE2<Op, std::void_t<decltype(Op::kUniformE2)>>
    : std::integral_constant<bool, Op::kUniformE2> {};
// Ops whose kUnmask is true may run a uniform tile without their pair
// mask when the call allows it (unmask_allowed) and, for ops that also
// declare self_safe, every j of the tile passes that (see pair_body).
template <class Op, class = void>
struct HasUnmask : std::false_type {};
template <class Op>
struct HasUnmask<Op, std::void_t<decltype(Op::kUnmask)>>
    : std::integral_constant<bool, Op::kUnmask> {};
template <class Op, class = void>
struct HasSelfSafe : std::false_type {};
template <class Op>
struct HasSelfSafe<Op, std::void_t<decltype(&Op::self_safe)>> : std::true_type {};
template <class Op, bool = HasUniformE2<Op>::value>
struct SoftType {
    using type = double;
};
template <class Op>
struct SoftType<Op, true> {
    using type = decltype(Op::slot_e2(std::declval<const typename Op::Slot&>()));
};

// FP64 fallback of the split-FP32 interactions.  An unsoftened close pair's
// terms (m / R^3, m^2 / R^6 ...) overflow FP32 long before FP64, and an
// overflow in any term reaches the sums as an infinity or a NaN (every pair
// term is a product with the overflowing factor), so a finite result needs
// no second look and the test costs a few instructions per i, not per pair.
// An i whose sums come out non-finite is evaluated again by the FP64
// interaction of the same kernel on the FP64 inputs:
//   Op::Fallback64      in the pair kernel's epilogue, over the same j-range
//                       (Acc, AccJerk, Tstep: their FP64 loops fit in the
//                       registers the FP32 loop already holds);
//   Op::PostFallback64  by fallback_kernel after the call, over every j, one
//                       thread per i (PNAcc, whose FP64 polynomial inlined
//                       in the pair kernel would cost it occupancy).
// Non-finite inputs give non-finite FP64 sums, as in the FP64 kernel.
template <class Op, class = void>
struct HasFallback : std::false_type {};
template <class Op>
struct HasFallback<Op, std::void_t<typename Op::Fallback64>> : std::true_type {};
template <class Op, class = void>
struct HasPostFallback : std::false_type {};
template <class Op>
struct HasPostFallback<Op, std::void_t<typename Op::PostFallback64>> : std::true_type {};
// Ops that combine two interactions check each one's sums and evaluate
// only a non-finite part again (fallback_parts(a, i, j0, j1, jstep, tot)),
// as the two separate calls would.
template <class Op, class = void>
struct HasPartFallback : std::false_type {};
template <class Op>
struct HasPartFallback<Op, std::void_t<decltype(&Op::fallback_parts)>> : std::true_type {};
// Ops whose per-particle store is spread over the lanes of the small-set
// kernel's group (store_group(a, i, totals, lane)) instead of one lane.
template <class Op, class = void>
struct HasGroupStore : std::false_type {};
template <class Op>
struct HasGroupStore<Op, std::void_t<decltype(&Op::store_group)>> : std::true_type {};

template <int N>
__device__ __forceinline__ bool all_finite(const double (&t)[N])
{
    bool ok = true;
#pragma unroll
    for (int k = 0; k < N; ++k) ok = ok && __builtin_isfinite(t[k]);
    return ok;
}

// tot = the sums of i over j = j0, j0 + jstep, ... below j1 by the FP64
// interaction F on the FP64 inputs.
template <class F>
__device__ void fallback64(const typename F::Args& a, int i, int j0, int j1, int jstep,
                           double* tot)
{
    const typename F::Frame fr = F::frame(a, 0);
    typename F::IReg r;
    typename F::Acc acc;
    F::load_i(a, fr, i, r);
    F::zero(acc);
    for (int j = j0; j < j1; j += jstep) {
        typename F::Slot s;
        F::load_j(a, fr, j, s);
        F::interact(r, s, acc);
    }
    F::flush(acc);
    for (int k = 0; k < F::kNacc; ++k) tot[k] = F::get(acc, k);
}

// A value every lane of the wave holds, moved to scalar registers.
__device__ __forceinline__ float uniform_value(float v)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_readfirstlane(__builtin_bit_cast(int, v)));
}
__device__ __forceinline__ double uniform_value(double v)
{
    const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
    const unsigned lo = static_cast<unsigned>(
        __builtin_amdgcn_readfirstlane(static_cast<int>(static_cast<unsigned>(b))));
    const unsigned hi = static_cast<unsigned>(
        __builtin_amdgcn_readfirstlane(static_cast<int>(static_cast<unsigned>(b >> 32))));
    return __builtin_bit_cast(double, (static_cast<unsigned long long>(hi) << 32) | lo);
}

// One pair of a tile pass: uniform (kU) or not, masked (kM) or not.
template <class Op, bool kU, bool kM, class E2>
__device__ __forceinline__ void tile_pair(const typename Op::IReg& r,
                                          const typename Op::Slot& s, typename Op::Acc& c,
                                          E2 e2u)
{
    if constexpr (kU && !kM) {
        Op::template interact<true, false>(r, s, c, e2u);
    } else if constexpr (kU) {
        Op::template interact<true>(r, s, c, e2u);
    } else {
        Op::interact(r, s, c);
    }
}

// One tile's interactions: UNROLL slots at a time, each applied to the WPT
// i-particles, accumulators flushed every CH slots.
template <class Op, int WPT, int UNROLL, int CH, bool kU, bool kM, class E2>
__device__ __forceinline__ void tile_pass(const typename Op::Slot* tile, int cnt,
                                          const typename Op::IReg (&ir)[WPT],
                                          typename Op::Acc (&acc)[WPT], E2 e2u)
{
    using Slot = typename Op::Slot;
    for (int k0 = 0; k0 < cnt; k0 += CH) {
        const int kend = min(cnt, k0 + CH);
        int k = k0;
        for (; k + UNROLL <= kend; k += UNROLL) {
#pragma unroll
            for (int u = 0; u < UNROLL; ++u) {
                const Slot s = tile[k + u];
#pragma unroll
                for (int w = 0; w < WPT; ++w) tile_pair<Op, kU, kM>(ir[w], s, acc[w], e2u);
            }
        }
        for (; k < kend; ++k) {
            const Slot s = tile[k];
#pragma unroll
            for (int w = 0; w < WPT; ++w) tile_pair<Op, kU, kM>(ir[w], s, acc[w], e2u);
        }
#pragma unroll
        for (int w = 0; w < WPT; ++w) Op::flush(acc[w]);
    }
}

// The tile engine's work for workgroup `bid` of a call (the body of
// pair_kernel; also driven by the device-scheduled block steps).
template <class Op, int BLOCK, int WPT, int TILE, int UNROLL>
__device__ __forceinline__ void pair_body(const typename Op::Args& a, int bid)
{
    using Slot = typename Op::Slot;
    using IReg = typename Op::IReg;
    using Acc = typename Op::Acc;
    constexpr bool kUni = HasUniformE2<Op>::value;
    __shared__ Slot tile[TILE];

    const GridPos g = grid_pos_of(bid, a.nsplit);
    const int tid = static_cast<int>(threadIdx.x);
    const int ibase = g.iblock * (BLOCK * WPT);
    const typename Op::Frame fr = Op::frame(a, ibase);

    IReg ir[WPT];
    Acc acc[WPT];
#pragma unroll
    for (int w = 0; w < WPT; ++w) {
        int i = ibase + w * BLOCK + tid;
        Op::load_i(a, fr, i < a.ni ? i : a.ni - 1, ir[w]);
        Op::zero(acc[w]);
    }

    // Uniform softening: e2 of the block's first i.  A tile takes the
    // uniform path when every i of the workgroup and every j of the tile
    // carry exactly that value; the pair value E + E is then formed once,
    // the same number the per-pair sum gives, so results do not change.
    // Ops with unmask_allowed run a uniform tile unmasked when the call
    // allows it (and, with self_safe, every j of the tile passes that):
    // their sums are then the masked ones.
    constexpr bool kUnmask = kUni && HasUnmask<Op>::value;
    constexpr bool kSelf = kUnmask && HasSelfSafe<Op>::value;
    bool iuni = true;
    [[maybe_unused]] typename SoftType<Op>::type e2_first{};
    [[maybe_unused]] bool unmask = false;
    [[maybe_unused]] typename SoftType<Op>::type y_self{};
    if constexpr (kUni) {
        e2_first = Op::e2_of_i(a, ibase);
#pragma unroll
        for (int w = 0; w < WPT; ++w) {
            const int i = ibase + w * BLOCK + tid;
            iuni = iuni && (i >= a.ni || Op::e2_of_i(a, i) == e2_first);
        }
    }
    if constexpr (kUnmask) unmask = Op::unmask_allowed(a);
    if constexpr (kSelf) {
        // the same for every thread: held in scalar registers, so the pair
        // loop keeps its vector registers
        y_self = uniform_value(Op::self_y(e2_first + e2_first));
    }

    const int j0 = g.jsplit * a.jchunk;
    const int j1 = min(a.nj, j0 + a.jchunk);
    constexpr int CH = Op::kFlush < TILE ? Op::kFlush : TILE;

    for (int jt = j0; jt < j1; jt += TILE) {
        const int cnt = min(TILE, j1 - jt);
        __syncthreads();
        bool tuni = iuni;
        [[maybe_unused]] bool tsafe = unmask;
        for (int k = tid; k < cnt; k += BLOCK) {
            Op::load_j(a, fr, jt + k, tile[k]);
            if constexpr (kUni) tuni = tuni && Op::slot_e2(tile[k]) == e2_first;
            if constexpr (kSelf) tsafe = tsafe && Op::self_safe(y_self, tile[k]);
        }
        if constexpr (kUnmask) {
            // a uniform tile that may not run unmasked takes the general
            // loop (the same sums; one more add per pair), so the kernel
            // holds two pair loops, not three
            if (__syncthreads_and(tuni && tsafe)) {
                tile_pass<Op, WPT, UNROLL, CH, true, false>(tile, cnt, ir, acc, e2_first + e2_first);
            } else {
                tile_pass<Op, WPT, UNROLL, CH, false, true>(tile, cnt, ir, acc, e2_first);
            }
        } else if constexpr (kUni) {
            if (__syncthreads_and(tuni)) {
                tile_pass<Op, WPT, UNROLL, CH, true, true>(tile, cnt, ir, acc, e2_first + e2_first);
            } else {
                tile_pass<Op, WPT, UNROLL, CH, false, true>(tile, cnt, ir, acc, e2_first);
            }
        } else {
            __syncthreads();
            tile_pass<Op, WPT, UNROLL, CH, false, true>(tile, cnt, ir, acc, 0);
        }
    }

#pragma unroll
    for (int w = 0; w < WPT; ++w) {
        const int i = ibase + w * BLOCK + tid;
        if (i >= a.ni) continue;
        double tot[Op::kNacc];
#pragma unroll
        for (int k = 0; k < Op::kNacc; ++k) tot[k] = Op::get(acc[w], k);
        if constexpr (HasFallback<Op>::value) {
            if (!all_finite(tot)) fallback64<typename Op::Fallback64>(a, i, j0, j1, 1, tot);
        }
        if (a.nsplit == 1) {
            Op::store(a, i, tot);
        } else {
            double* part = a.partial + static_cast<size_t>(g.jsplit) * Op::kNacc * a.ni;
#pragma unroll
            for (int k = 0; k < Op::kNacc; ++k) part[static_cast<size_t>(k) * a.ni + i] = tot[k];
        }
    }
}

template <class Op, int BLOCK, int WPT, int TILE, int UNROLL>
__global__ __launch_bounds__(BLOCK) void pair_kernel(typename Op::Args a)
{
    pair_body<Op, BLOCK, WPT, TILE, UNROLL>(a, static_cast<int>(blockIdx.x));
}

// Small sets in one launch: G consecutive lanes (G divides 64) share an
// i-particle, BLOCK / G of them per workgroup; of every LDS tile lane group
// member g takes the slots k = g + G m (interleaved, so the group reads
// neighbouring slots); the G partial sums of an i are then combined in a
// fixed shuffle tree (lanes 2k and 2k + 1, then 4k and 4k + 2, ...), so
// results are the same on every run and the call needs no j-split partials
// or second kernel.
template <class Op, int BLOCK, int G, int TILE>
__device__ __forceinline__ void pair_small_body(const typename Op::Args& a, int bid)
{
    using Slot = typename Op::Slot;
    static_assert(64 % G == 0 && TILE % G == 0 && BLOCK % 64 == 0,
                  "G must divide 64 and TILE; BLOCK whole waves");
    constexpr int kIpb = BLOCK / G;
    __shared__ Slot tile[TILE];
    const int tid = static_cast<int>(threadIdx.x);
    const int il = tid / G;
    const int g = tid % G;
    const int ibase = bid * kIpb;
    const typename Op::Frame fr = Op::frame(a, ibase);
    const int i = ibase + il;
    typename Op::IReg ir;
    typename Op::Acc acc;
    Op::load_i(a, fr, i < a.ni ? i : a.ni - 1, ir);
    Op::zero(acc);
    for (int jt = 0; jt < a.nj; jt += TILE) {
        const int cnt = min(TILE, a.nj - jt);
        __syncthreads();
        for (int k = tid; k < cnt; k += BLOCK) Op::load_j(a, fr, jt + k, tile[k]);
        __syncthreads();
        // TILE / G slots per lane per tile, within any Op's flush interval
        for (int k = g; k < cnt; k += G) Op::interact(ir, tile[k], acc);
        Op::flush(acc);
    }
    double tot[Op::kNacc];
#pragma unroll
    for (int k = 0; k < Op::kNacc; ++k) tot[k] = Op::get(acc, k);
    if constexpr (HasPartFallback<Op>::value) {
        // lane g took the j with j % G == g (TILE is a multiple of G)
        if (i < a.ni) Op::fallback_parts(a, i, g, a.nj, G, tot);
    } else if constexpr (HasFallback<Op>::value) {
        // lane g took the j with j % G == g (TILE is a multiple of G)
        if (i < a.ni && !all_finite(tot))
            fallback64<typename Op::Fallback64>(a, i, g, a.nj, G, tot);
    }
#pragma unroll
    for (int off = 1; off < G; off <<= 1) {
        double other[Op::kNacc];
#pragma unroll
        for (int k = 0; k < Op::kNacc; ++k) other[k] = __shfl_xor(tot[k], off);
        if ((g & (2 * off - 1)) == 0) Op::combine(tot, other);
    }
    if constexpr (HasGroupStore<Op>::value) {
        // the group's totals (lane 0) to every lane of the group
#pragma unroll
        for (int k = 0; k < Op::kNacc; ++k) tot[k] = __shfl(tot[k], 0, G);
        if (i < a.ni) Op::store_group(a, i, tot, g);
    } else {
        if (g == 0 && i < a.ni) Op::store(a, i, tot);
    }
}

template <class Op, int BLOCK, int G, int TILE>
__global__ __launch_bounds__(BLOCK) void pair_small_kernel(typename Op::Args a)
{
    pair_small_body<Op, BLOCK, G, TILE>(a, static_cast<int>(blockIdx.x));
}

// Combine the j-split partials of each i in split order and store.  The
// partials are loaded kBatch splits at a time before they are combined, so a
// thread has kBatch * kNacc loads in flight instead of kNacc: for small ni
// (a few waves) the kernel is bound by memory latency, not bandwidth.
template <class Op>
__device__ __forceinline__ void reduce_body(const typename Op::Args& a, int bid)
{
    constexpr int kBatch = 8;
    const int i = bid * 256 + static_cast<int>(threadIdx.x);
    if (i >= a.ni) return;
    double tot[Op::kNacc];
    const size_t stride = static_cast<size_t>(Op::kNacc) * a.ni;
#pragma unroll
    for (int k = 0; k < Op::kNacc; ++k) tot[k] = a.partial[static_cast<size_t>(k) * a.ni + i];
    for (int s0 = 1; s0 < a.nsplit; s0 += kBatch) {
        double part[kBatch][Op::kNacc];
#pragma unroll
        for (int b = 0; b < kBatch; ++b) {
            if (s0 + b < a.nsplit) {
                const double* src = a.partial + static_cast<size_t>(s0 + b) * stride + i;
#pragma unroll
                for (int k = 0; k < Op::kNacc; ++k) part[b][k] = src[static_cast<size_t>(k) * a.ni];
            }
        }
#pragma unroll
        for (int b = 0; b < kBatch; ++b)
            if (s0 + b < a.nsplit) Op::combine(tot, part[b]);
    }
    Op::store(a, i, tot);
}

template <class Op>
__global__ __launch_bounds__(256) void reduce_kernel(typename Op::Args a)
{
    reduce_body<Op>(a, static_cast<int>(blockIdx.x));
}

// The same combine for small ni, where one thread per i leaves most of the
// chip idle and waits on memory latency once per batch: T consecutive lanes
// share an i, lane t combining splits [t c, (t + 1) c) in order (c = ceil(
// nsplit / T)), then the T lane totals are combined pairwise in a fixed tree
// (lanes 2k and 2k + 1, then 4k and 4k + 2, ...), so the result is the same
// on every run.
template <class Op, int T>
__device__ __forceinline__ void reduce_wide_body(const typename Op::Args& a, int bid)
{
    constexpr int kBatch = 8;
    static_assert(T > 1 && (T & (T - 1)) == 0 && T <= 64, "T must be a power of two <= 64");
    const int gid = bid * 256 + static_cast<int>(threadIdx.x);
    const int i = gid / T;
    const int t = gid % T;
    const bool valid = i < a.ni;
    const int c = (a.nsplit + T - 1) / T;
    const int sb = t * c;
    const int se = min(a.nsplit, sb + c);
    const size_t stride = static_cast<size_t>(Op::kNacc) * a.ni;
    double tot[Op::kNacc];
#pragma unroll
    for (int k = 0; k < Op::kNacc; ++k) tot[k] = 0.0;
    int have = 0;
    for (int s0 = sb; valid && s0 < se; s0 += kBatch) {
        double part[kBatch][Op::kNacc];
#pragma unroll
        for (int b = 0; b < kBatch; ++b) {
            if (s0 + b < se) {
                const double* src = a.partial + static_cast<size_t>(s0 + b) * stride + i;
#pragma unroll
                for (int k = 0; k < Op::kNacc; ++k) part[b][k] = src[static_cast<size_t>(k) * a.ni];
            }
        }
#pragma unroll
        for (int b = 0; b < kBatch; ++b) {
            if (s0 + b < se) {
                if (have) {
                    Op::combine(tot, part[b]);
                } else {
#pragma unroll
                    for (int k = 0; k < Op::kNacc; ++k) tot[k] = part[b][k];
                    have = 1;
                }
            }
        }
    }
    // every lane of the wave takes part in the exchanges
#pragma unroll
    for (int off = 1; off < T; off <<= 1) {
        double other[Op::kNacc];
#pragma unroll
        for (int k = 0; k < Op::kNacc; ++k) other[k] = __shfl_xor(tot[k], off);
        const int ohave = __shfl_xor(have, off);
        if ((t & (2 * off - 1)) == 0 && ohave) {
            if (have) {
                Op::combine(tot, other);
            } else {
#pragma unroll
                for (int k = 0; k < Op::kNacc; ++k) tot[k] = other[k];
                have = 1;
            }
        }
    }
    if (valid && t == 0) Op::store(a, i, tot);
}

template <class Op, int T>
__global__ __launch_bounds__(256) void reduce_wide_kernel(typename Op::Args a)
{
    reduce_wide_body<Op, T>(a, static_cast<int>(blockIdx.x));
}

// Op::PostFallback64 after a call: every i whose stored outputs (the sums,
// for the ops that have one) are non-finite, again in FP64 over every j.
template <class Op>
__global__ __launch_bounds__(256) void fallback_kernel(typename Op::Args a)
{
    using F = typename Op::PostFallback64;
    static_assert(F::kNacc == Op::kNacc, "fallback must produce the same sums");
    const int i = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= a.ni) return;
    double tot[Op::kNacc];
#pragma unroll
    for (int k = 0; k < Op::kNacc; ++k) tot[k] = static_cast<double>(a.out[k][i]);
    if (all_finite(tot)) return;
    fallback64<F>(a, i, 0, a.nj, 1, tot);
    Op::store(a, i, tot);
}

// Launch constants shared by the host launcher (pairwise.hip) and the
// device-scheduled block steps (sia.hip), which must plan every call alike.

// Tile shapes {BLOCK, WPT, TILE, UNROLL} of the light interactions (Phi,
// Acc, AccJerk, Tstep, split forms); the last, one wave per workgroup, is
// for small systems.  The tuning table picks one per kernel, mode and size.
constexpr int kNumShapes = 6;
constexpr int kLightShapeDims[kNumShapes][4] = {
    {256, 2, 256, 4}, {512, 2, 512, 2}, {256, 4, 256, 2}, {128, 2, 128, 4}, {512, 1, 512, 4},
    {64, 1, 64, 4}};
// Below this many i-particles the j-split combine uses 8 lanes per i.
constexpr int kWideReduceMaxNi = 32768;

// Calls of at most kSmallPairs pairs with kSmallMinJ < nj <= kSmallMaxJ
// (in any mode but ieee64, whose sums keep the reference's order) take
// pair_small_kernel: a few-thousand-particle system in one launch instead of
// j-split partials and a combine kernel.  Sets of at most kSmallMinJ stay on
// the tile engine, whose one-tile sums the device SIA subtrees reproduce bit
// for bit; a few i against many j (block-step levels) keep the j-splits,
// which spread j over the chip.
constexpr long long kSmallPairs = 1ll << 21;
constexpr int kSmallMinJ = 256;
constexpr int kSmallMaxJ = 2048;
constexpr int kSmallBlock = 256;
constexpr int kSmallGroups = 64;
// the largest power-of-two tile of at most 1024 slots within 64 KiB of LDS
// (one tile, and one wait on global memory, for a 1024-particle set)
template <class Op>
constexpr int small_tile()
{
    int t = 1024;
    while (t > kSmallGroups && t * static_cast<int>(sizeof(typename Op::Slot)) > 65536) t /= 2;
    return t;
}

// Launch geometry of one all-pairs call.
struct Launch {
    int nblock_i;  // i blocks
    int nsplit;    // j splits
    int jchunk;    // j per split (multiple of the minimum chunk)
};

// Pick the j-split so that the grid has at least `target_wg` workgroups
// (several per CU) without cutting j ranges below `min_chunk` j-particles
// (a multiple of 16; the tile loop handles a partly filled last tile).
// Smallest j-chunk of a split: whole tiles for a set that fits one tile
// (such a set is always one split, so its sums are the one-tile order) and,
// in the reference-order mode, never below 256 particles; 32 particles
// otherwise, so that a set of a few tiles still spreads over the chip.
__host__ __device__ inline int min_jchunk(int nj, int tile, bool exact)
{
    if (exact) return tile > 256 ? tile : 256;
    return nj > tile ? 32 : tile;
}

__host__ __device__ inline Launch plan_launch(int ni, int nj, int block, int wpt, int tile,
                                              int target_wg, int max_split, int min_chunk)
{
    Launch L;
    const int per_block = block * wpt;
    L.nblock_i = (ni + per_block - 1) / per_block;
    int want = (target_wg + L.nblock_i - 1) / L.nblock_i;
    int max_by_j = (nj + min_chunk - 1) / min_chunk;
    int ns = want < max_by_j ? want : max_by_j;
    if (ns > max_split) ns = max_split;
    if (ns < 1) ns = 1;
    // j range per split, rounded to whole multiples of min_chunk, then the
    // split count that range actually needs
    int chunk = (nj + ns - 1) / ns;
    chunk = ((chunk + min_chunk - 1) / min_chunk) * min_chunk;
    if (chunk < min_chunk) chunk = min_chunk;
    L.jchunk = chunk;
    L.nsplit = nj > 0 ? (nj + chunk - 1) / chunk : 1;
    return L;
}

}  // namespace tupan
