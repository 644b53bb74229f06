// Kepler-guess network: the packed BF16 inference block and one wave's BF16
// inference of 64 items, shared by the stand-alone inference kernel
// (mlp.hip) and the Sakura near pass that evaluates the guess in-line
// (sakura.hip), so both give the same guesses bit for bit.
//
// The network is a 4 -> 64 -> 64 -> 1 perceptron with SiLU activations that
// maps the dimensionless description of a universal-variable solve
// (kepler.h: SolveFeatures) to delta = s r / dt - 1.  Inputs are prepared
// as x = (log tau, sigma, kappa, eps), standardised by mean and inv_std; the
// output y maps to delta = y * out_scale + out_shift.
#pragma once

#include <hip/hip_runtime.h>

namespace tupan {

constexpr int kHidden = 64;
constexpr int kIn = 4;

// Packed BF16 inference block (tupan_mlp_pack).
struct NetPacked {
    float mean[4];
    float inv_std[4];
    float out_scale, out_shift;
    float pad0[6];
    float b1[kHidden];
    float b2[kHidden];
    float b3;
    float pad1[15];
    unsigned short w1[kHidden * kIn];
    unsigned short w2[kHidden * kHidden];
    unsigned short w3[kHidden];
};

// Round to nearest even BF16 (finite inputs; NaN stays NaN).
__host__ __device__ __forceinline__ unsigned short to_bf16(float x)
{
    unsigned int u = __builtin_bit_cast(unsigned int, x);
    if ((u & 0x7f800000u) == 0x7f800000u) return static_cast<unsigned short>(u >> 16 | ((u & 0xffffu) ? 0x40u : 0u));
    u += 0x7fffu + ((u >> 16) & 1u);
    return static_cast<unsigned short>(u >> 16);
}

__device__ __forceinline__ float from_bf16(unsigned short h)
{
    return __builtin_bit_cast(float, static_cast<unsigned int>(h) << 16);
}

__device__ __forceinline__ float sigmoid(float z) { return 1.0f / (1.0f + __expf(-z)); }

// Inference forms: the hardware's round-to-nearest-even BF16 conversion
// (v_cvt_pk_bf16_f32; the same value as to_bf16 for every finite input) and
// SiLU with the hardware reciprocal (1 ulp) in place of an IEEE division.
__device__ __forceinline__ unsigned short bf16_rne(float x)
{
    return __builtin_bit_cast(unsigned short, static_cast<__bf16>(x));
}

__device__ __forceinline__ float silu_infer(float z)
{
    return z * __builtin_amdgcn_rcpf(1.0f + __expf(-z));
}

// x = (log tau, sigma, kappa, eps), standardised.
__device__ __forceinline__ void prepare_features(float tau, float sigma, float kappa, float eps,
                                                 const float* mean, const float* inv_std,
                                                 float x[4])
{
    x[0] = __logf(fmaxf(tau, 1e-30f));
    x[1] = sigma;
    x[2] = kappa;
    x[3] = eps;
#pragma unroll
    for (int c = 0; c < 4; ++c) x[c] = (x[c] - mean[c]) * inv_std[c];
}

// ---------------------------------------------------------------------------
// One wave, 64 items: lanes l and l + 32 share item (l mod 32) (+ 32 for the
// second column tile) and each computes half of its 64 layer-1 activations,
// exactly the layer-2 B operand it supplies:
//   v_mfma_f32_32x32x8_bf16:  A[m][k] from lane m + 32 (k / 4), element k % 4
//                             B[k][n] from lane n + 32 (k / 4), element k % 4
//                             D[m][n] in lane n + 32 ((m / 4) % 2),
//                                     element (m % 4) + 4 (m / 8)
// Layer 2 is D = W2 (64 x 64, hidden x k) times H1^T (64 x 64, k x item):
// 2 x 2 output tiles of 32 x 32, 8 k-steps each.  D's column is the item,
// so layer 3 (a dot over the hidden rows) sums within the lane and across
// the lane pair (l, l ^ 32).
// ---------------------------------------------------------------------------
typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

// Copy of the packed block into LDS by the whole workgroup (then a
// barrier); `src` 16-byte aligned.
__device__ __forceinline__ void mlp_stage(const NetPacked* src, NetPacked* dst)
{
    static_assert(sizeof(NetPacked) % 16 == 0, "the network block is copied in 16-byte units");
    const uint4* s = reinterpret_cast<const uint4*>(src);
    uint4* d = reinterpret_cast<uint4*>(dst);
    for (int k = static_cast<int>(threadIdx.x); k < static_cast<int>(sizeof(NetPacked) / 16);
         k += static_cast<int>(blockDim.x))
        d[k] = s[k];
    __syncthreads();
}

// This lane's A operand (W2) of row tile r, k-step s: W2[m][k0 .. k0 + 3]
// with m = 32 r + lane mod 32, k0 = 8 s + 4 (lane / 32), one 8-byte read.
__device__ __forceinline__ bf16x4 mlp_w2_operand(const NetPacked* net, int r, int s, int lane)
{
    const int m = 32 * r + (lane & 31);
    const int k0 = 8 * s + 4 * ((lane & 63) >> 5);
    return *reinterpret_cast<const bf16x4*>(&net->w2[m * kHidden + k0]);
}

// x[c]: the prepared inputs of item 32 c + (lane mod 32).  Returns in y[c]
// the network output (before b3 and the output scaling) of that item, the
// same in lanes l and l ^ 32.  Every lane of the wave must take part.  Each
// weight is read once and applied to both column tiles, and the W2 operands
// are read as they are used, so few weights are live at a time.
__device__ __forceinline__ void mlp_wave_eval(const NetPacked* net, const float x[2][4], int lane,
                                              float y[2])
{
    const int half = (lane & 63) >> 5;
    float xb[2][4];
#pragma unroll
    for (int c = 0; c < 2; ++c) {
#pragma unroll
        for (int q = 0; q < 4; ++q) xb[c][q] = from_bf16(bf16_rne(x[c][q]));
    }
    bf16x4 B[2][8];
#pragma unroll
    for (int s = 0; s < 8; ++s) {
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int k = 8 * s + 4 * half + e;
            const float b = net->b1[k];
            float w[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) w[q] = from_bf16(net->w1[k * kIn + q]);
#pragma unroll
            for (int c = 0; c < 2; ++c) {
                float z = b;
#pragma unroll
                for (int q = 0; q < 4; ++q) z = fmaf(w[q], xb[c][q], z);
                B[c][s][e] = static_cast<short>(bf16_rne(silu_infer(z)));
            }
        }
    }
    y[0] = y[1] = 0.f;
#pragma unroll
    for (int r = 0; r < 2; ++r) {
        f32x16 acc[2];
#pragma unroll
        for (int c = 0; c < 2; ++c) {
#pragma unroll
            for (int v = 0; v < 16; ++v) acc[c][v] = 0.f;
        }
#pragma unroll
        for (int s = 0; s < 8; ++s) {
            const bf16x4 a = mlp_w2_operand(net, r, s, lane);
#pragma unroll
            for (int c = 0; c < 2; ++c)
                acc[c] = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, B[c][s], acc[c], 0, 0, 0);
        }
#pragma unroll
        for (int v = 0; v < 16; ++v) {
            const int m = 32 * r + (v % 4) + 8 * (v / 4) + 4 * half;
            const float b = net->b2[m];
            const float w = from_bf16(net->w3[m]);
#pragma unroll
            for (int c = 0; c < 2; ++c) {
                const float z = acc[c][v] + b;
                const float h = from_bf16(bf16_rne(silu_infer(z)));
                y[c] = fmaf(w, h, y[c]);
            }
        }
    }
#pragma unroll
    for (int c = 0; c < 2; ++c) y[c] += __shfl_xor(y[c], 32, 64);
}

// delta from the network output of mlp_wave_eval.
__device__ __forceinline__ float mlp_delta(const NetPacked* net, float y)
{
    return (y + net->b3) * net->out_scale + net->out_shift;
}

}  // namespace tupan
