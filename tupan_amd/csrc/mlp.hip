// Kepler-guess network for gfx950: a 4 -> 64 -> 64 -> 1 perceptron with
// SiLU activations that maps the dimensionless description of a universal-
// variable solve (kepler.h: SolveFeatures) to delta = s r / dt - 1, the
// relative distance of the converged anomaly s from the plain guess dt / r.
//
// Parameters (FP32 master copy, flat, kNetParams floats):
//   W1[64][4] b1[64] W2[64][64] b2[64] W3[64] b3
// Normalisation (kNormFloats floats): mean[4], inv_std[4] of the prepared
// inputs x = (log tau, sigma, kappa, eps), and out_scale, out_shift: the
// network output y maps to delta = y * out_scale + out_shift, and training
// fits y to (delta - out_shift) / out_scale.
//
// Kernels:
//   mlp_grad_kernel / mlp_grad_reduce_kernel  FP32 forward + backward of the
//       mean squared error over a batch; per-block partial gradients are
//       summed in block order (deterministic)
//   mlp_forward_kernel   FP32 forward (evaluation)
//   mlp_adam_kernel      Adam update
//   mlp_pack_kernel      BF16 inference block from the FP32 parameters
//   mlp_infer_kernel     BF16 inference with FP32 accumulation; the 64x64
//       layer runs on the matrix cores (v_mfma_f32_32x32x8_bf16), the two
//       thin layers on the vector ALUs
#include "mlp_wave.h"
#include "tupan_internal.h"

namespace tupan {

constexpr int kOffW1 = 0;
constexpr int kOffB1 = kOffW1 + kHidden * kIn;       // 256
constexpr int kOffW2 = kOffB1 + kHidden;             // 320
constexpr int kOffB2 = kOffW2 + kHidden * kHidden;   // 4416
constexpr int kOffW3 = kOffB2 + kHidden;             // 4480
constexpr int kOffB3 = kOffW3 + kHidden;             // 4544
constexpr int kNetParams = kOffB3 + 1;               // 4545
constexpr int kNormFloats = 10;

// x = (log tau, sigma, kappa, eps) of item t of the SoA features, standardised.
__device__ __forceinline__ void prepare(const float* feat, size_t n, size_t t,
                                        const float* mean, const float* inv_std, float x[4])
{
    prepare_features(feat[0 * n + t], feat[1 * n + t], feat[2 * n + t], feat[3 * n + t], mean,
                     inv_std, x);
}

// ---------------------------------------------------------------------------
// Training: FP32 forward + backward.  One sample per thread, kBatchBlock
// samples per workgroup; the parameters, the hidden activations and the
// layer-2 deltas of the workgroup's samples sit in LDS.
// ---------------------------------------------------------------------------
constexpr int kBatchBlock = 128;

__global__ __launch_bounds__(kBatchBlock) void mlp_grad_kernel(int n, const float* feat,
                                                               const float* target,
                                                               const float* params,
                                                               const float* norm,
                                                               float* gpart, float* lpart)
{
    __shared__ float P[kNetParams];
    __shared__ float H1[kBatchBlock][kHidden + 1];
    __shared__ float D2[kBatchBlock][kHidden + 1];  // z2, then dz2, then dz1
    __shared__ float H2[kBatchBlock][kHidden + 1];
    __shared__ float X[kBatchBlock][4];
    __shared__ float DY[kBatchBlock];
    __shared__ float LS[kBatchBlock];
    const int tid = static_cast<int>(threadIdx.x);
    for (int k = tid; k < kNetParams; k += kBatchBlock) P[k] = params[k];
    __syncthreads();
    const size_t t = static_cast<size_t>(blockIdx.x) * kBatchBlock + tid;
    const bool valid = t < static_cast<size_t>(n);
    const float* mean = norm;
    const float* inv_std = norm + 4;
    const float out_scale = norm[8];
    const float out_shift = norm[9];
    float x[4] = {0.f, 0.f, 0.f, 0.f};
    if (valid) prepare(feat, static_cast<size_t>(n), t, mean, inv_std, x);
#pragma unroll
    for (int c = 0; c < 4; ++c) X[tid][c] = x[c];

    // forward
    float z1[kHidden];
#pragma unroll
    for (int k = 0; k < kHidden; ++k) {
        float z = P[kOffB1 + k];
#pragma unroll
        for (int c = 0; c < 4; ++c) z = fmaf(P[kOffW1 + k * kIn + c], x[c], z);
        z1[k] = z;
        H1[tid][k] = z * sigmoid(z);
    }
    float y = P[kOffB3];
    for (int m = 0; m < kHidden; ++m) {
        float z = P[kOffB2 + m];
#pragma unroll 8
        for (int k = 0; k < kHidden; ++k) z = fmaf(P[kOffW2 + m * kHidden + k], H1[tid][k], z);
        const float h = z * sigmoid(z);
        D2[tid][m] = z;
        H2[tid][m] = h;
        y = fmaf(P[kOffW3 + m], h, y);
    }
    const float tn = valid ? (target[t] - out_shift) / out_scale : 0.f;
    const float r = valid ? y - tn : 0.f;
    LS[tid] = r * r;
    const float dy = 2.0f * r / static_cast<float>(n);
    DY[tid] = dy;
    // backward through layer 3 and the second SiLU
    for (int m = 0; m < kHidden; ++m) {
        const float z = D2[tid][m];
        const float sg = sigmoid(z);
        D2[tid][m] = dy * P[kOffW3 + m] * (sg * (1.0f + z * (1.0f - sg)));
    }
    // dz1 = (W2^T dz2) * silu'(z1), kept in registers until D2 is free
    float dz1[kHidden];
#pragma unroll
    for (int k = 0; k < kHidden; ++k) dz1[k] = 0.f;
    for (int m = 0; m < kHidden; ++m) {
        const float d = D2[tid][m];
#pragma unroll
        for (int k = 0; k < kHidden; ++k) dz1[k] = fmaf(P[kOffW2 + m * kHidden + k], d, dz1[k]);
    }
#pragma unroll
    for (int k = 0; k < kHidden; ++k) {
        const float sg = sigmoid(z1[k]);
        dz1[k] *= sg * (1.0f + z1[k] * (1.0f - sg));
    }
    __syncthreads();

    // workgroup sums of the layer-2 and layer-3 gradients
    float* g = gpart + static_cast<size_t>(blockIdx.x) * kNetParams;
    for (int q = tid; q < kHidden * kHidden; q += kBatchBlock) {
        const int m = q / kHidden, k = q % kHidden;
        float s = 0.f;
        for (int b = 0; b < kBatchBlock; ++b) s = fmaf(D2[b][m], H1[b][k], s);
        g[kOffW2 + q] = s;
    }
    if (tid < kHidden) {
        float sb = 0.f, sw = 0.f;
        for (int b = 0; b < kBatchBlock; ++b) {
            sb += D2[b][tid];
            sw = fmaf(DY[b], H2[b][tid], sw);
        }
        g[kOffB2 + tid] = sb;
        g[kOffW3 + tid] = sw;
    }
    if (tid == 0) {
        float sb = 0.f, sl = 0.f;
        for (int b = 0; b < kBatchBlock; ++b) {
            sb += DY[b];
            sl += LS[b];
        }
        g[kOffB3] = sb;
        lpart[blockIdx.x] = sl;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kHidden; ++k) D2[tid][k] = dz1[k];
    __syncthreads();
    for (int q = tid; q < kHidden * kIn; q += kBatchBlock) {
        const int k = q / kIn, c = q % kIn;
        float s = 0.f;
        for (int b = 0; b < kBatchBlock; ++b) s = fmaf(D2[b][k], X[b][c], s);
        g[kOffW1 + q] = s;
    }
    if (tid < kHidden) {
        float s = 0.f;
        for (int b = 0; b < kBatchBlock; ++b) s += D2[b][tid];
        g[kOffB1 + tid] = s;
    }
}

// Sum of the per-workgroup partials in workgroup order; loss = mean square.
__global__ __launch_bounds__(256) void mlp_grad_reduce_kernel(int nblocks, int n,
                                                              const float* gpart,
                                                              const float* lpart, float* grad,
                                                              float* loss)
{
    const int p = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    if (p < kNetParams) {
        float s = 0.f;
        for (int b = 0; b < nblocks; ++b) s += gpart[static_cast<size_t>(b) * kNetParams + p];
        grad[p] = s;
    }
    if (p == kNetParams && loss) {
        float s = 0.f;
        for (int b = 0; b < nblocks; ++b) s += lpart[b];
        loss[0] = s / static_cast<float>(n);
    }
}

__global__ __launch_bounds__(256) void mlp_forward_kernel(int n, const float* feat,
                                                          const float* params, const float* norm,
                                                          float* delta)
{
    __shared__ float P[kNetParams];
    for (int k = threadIdx.x; k < kNetParams; k += blockDim.x) P[k] = params[k];
    __syncthreads();
    const size_t t = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= static_cast<size_t>(n)) return;
    float x[4];
    prepare(feat, static_cast<size_t>(n), t, norm, norm + 4, x);
    float h1[kHidden];
#pragma unroll
    for (int k = 0; k < kHidden; ++k) {
        float z = P[kOffB1 + k];
#pragma unroll
        for (int c = 0; c < 4; ++c) z = fmaf(P[kOffW1 + k * kIn + c], x[c], z);
        h1[k] = z * sigmoid(z);
    }
    float y = P[kOffB3];
    for (int m = 0; m < kHidden; ++m) {
        float z = P[kOffB2 + m];
#pragma unroll
        for (int k = 0; k < kHidden; ++k) z = fmaf(P[kOffW2 + m * kHidden + k], h1[k], z);
        y = fmaf(P[kOffW3 + m], z * sigmoid(z), y);
    }
    delta[t] = y * norm[8] + norm[9];
}

__global__ __launch_bounds__(256) void mlp_adam_kernel(float* params, const float* grad,
                                                       float* m1, float* m2, float lr,
                                                       float beta1, float beta2, float eps,
                                                       float corr1, float corr2)
{
    const int p = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    if (p >= kNetParams) return;
    const float g = grad[p];
    const float a = beta1 * m1[p] + (1.0f - beta1) * g;
    const float b = beta2 * m2[p] + (1.0f - beta2) * g * g;
    m1[p] = a;
    m2[p] = b;
    params[p] -= lr * (a / corr1) / (sqrtf(b / corr2) + eps);
}

__global__ __launch_bounds__(256) void mlp_pack_kernel(const float* params, const float* norm,
                                                       NetPacked* out)
{
    const int p = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    if (p < 4) {
        out->mean[p] = norm[p];
        out->inv_std[p] = norm[4 + p];
    }
    if (p == 0) {
        out->out_scale = norm[8];
        out->out_shift = norm[9];
        out->b3 = params[kOffB3];
    }
    if (p < kHidden) {
        out->b1[p] = params[kOffB1 + p];
        out->b2[p] = params[kOffB2 + p];
        out->w3[p] = to_bf16(params[kOffW3 + p]);
    }
    if (p < kHidden * kIn) out->w1[p] = to_bf16(params[kOffW1 + p]);
    if (p < kHidden * kHidden) out->w2[p] = to_bf16(params[kOffW2 + p]);
}

// ---------------------------------------------------------------------------
// BF16 inference: one wave per 64 items (mlp_wave.h: mlp_wave_eval), the
// packed network staged in LDS.  No loop over item groups: the weights are
// then read where they are used instead of being hoisted into registers.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void mlp_infer_kernel(int n, const float* feat,
                                                        const NetPacked* gnet, float* delta)
{
    __shared__ __align__(16) NetPacked lnet;
    mlp_stage(gnet, &lnet);
    const NetPacked* net = &lnet;
    const int lane = static_cast<int>(threadIdx.x) & 63;
    const int half = lane >> 5;
    const int col = lane & 31;
    const size_t nn = static_cast<size_t>(n);
    const size_t grp = static_cast<size_t>(blockIdx.x) * 4 + (threadIdx.x >> 6);
    float x[2][4];
#pragma unroll
    for (int c = 0; c < 2; ++c) {
        const size_t item = grp * 64 + 32 * c + col;
#pragma unroll
        for (int q = 0; q < 4; ++q) x[c][q] = 0.f;
        if (item < nn) prepare(feat, nn, item, net->mean, net->inv_std, x[c]);
    }
    float y[2];
    mlp_wave_eval(net, x, lane, y);
    if (half == 0) {
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            const size_t item = grp * 64 + 32 * c + col;
            if (item < nn) delta[item] = mlp_delta(net, y[c]);
        }
    }
}

int mlp_infer(int n, const float* feat, const void* net, float* delta, hipStream_t stream)
{
    if (n <= 0) return TUPAN_OK;
    if (!feat || !net || !delta) return set_error(TUPAN_ERR_ARG, "null array");
    if (reinterpret_cast<uintptr_t>(net) % 16 != 0)
        return set_error(TUPAN_ERR_ARG, "the packed network must be 16-byte aligned");
    const int groups = (n + 63) / 64;
    const int grid = (groups + 3) / 4;
    hipLaunchKernelGGL(mlp_infer_kernel, dim3(grid), dim3(256), 0, stream, n, feat,
                       static_cast<const NetPacked*>(net), delta);
    return check_launch("mlp_infer");
}

}  // namespace tupan

using namespace tupan;

extern "C" {

int tupan_mlp_sizes(int* nparams, int* nnorm, int* packed_bytes)
{
    if (nparams) *nparams = kNetParams;
    if (nnorm) *nnorm = kNormFloats;
    if (packed_bytes) *packed_bytes = static_cast<int>(sizeof(NetPacked));
    return TUPAN_OK;
}

int tupan_mlp_grad(int n, const float* feat, const float* target, const float* params,
                   const float* norm, float* grad, float* loss, void* stream)
{
    if (n <= 0) return set_error(TUPAN_ERR_ARG, "empty batch");
    if (!feat || !target || !params || !norm || !grad) return set_error(TUPAN_ERR_ARG, "null array");
    hipStream_t s = static_cast<hipStream_t>(stream);
    const int nblocks = (n + kBatchBlock - 1) / kBatchBlock;
    float* gpart = static_cast<float*>(
        workspace((static_cast<size_t>(nblocks) * kNetParams + nblocks) * sizeof(float), s, 3));
    if (!gpart) return set_error(TUPAN_ERR_ALLOC, "workspace allocation failed");
    float* lpart = gpart + static_cast<size_t>(nblocks) * kNetParams;
    hipLaunchKernelGGL(mlp_grad_kernel, dim3(nblocks), dim3(kBatchBlock), 0, s, n, feat, target,
                       params, norm, gpart, lpart);
    hipLaunchKernelGGL(mlp_grad_reduce_kernel, dim3((kNetParams + 1 + 255) / 256), dim3(256), 0,
                       s, nblocks, n, gpart, lpart, grad, loss);
    return check_launch("mlp_grad");
}

int tupan_mlp_forward(int n, const float* feat, const float* params, const float* norm,
                      float* delta, void* stream)
{
    if (n <= 0) return TUPAN_OK;
    if (!feat || !params || !norm || !delta) return set_error(TUPAN_ERR_ARG, "null array");
    hipLaunchKernelGGL(mlp_forward_kernel, dim3((n + 255) / 256), dim3(256), 0,
                       static_cast<hipStream_t>(stream), n, feat, params, norm, delta);
    return check_launch("mlp_forward");
}

int tupan_mlp_adam(float* params, const float* grad, float* m1, float* m2, float lr,
                   float beta1, float beta2, float eps, int step, void* stream)
{
    if (!params || !grad || !m1 || !m2 || step < 1) return set_error(TUPAN_ERR_ARG, "bad Adam arguments");
    const float c1 = 1.0f - powf(beta1, static_cast<float>(step));
    const float c2 = 1.0f - powf(beta2, static_cast<float>(step));
    hipLaunchKernelGGL(mlp_adam_kernel, dim3((kNetParams + 255) / 256), dim3(256), 0,
                       static_cast<hipStream_t>(stream), params, grad, m1, m2, lr, beta1, beta2,
                       eps, c1, c2);
    return check_launch("mlp_adam");
}

int tupan_mlp_pack(const float* params, const float* norm, void* packed, void* stream)
{
    if (!params || !norm || !packed) return set_error(TUPAN_ERR_ARG, "null array");
    hipLaunchKernelGGL(mlp_pack_kernel, dim3((kHidden * kHidden + 255) / 256), dim3(256), 0,
                       static_cast<hipStream_t>(stream), params, norm,
                       static_cast<NetPacked*>(packed));
    return check_launch("mlp_pack");
}

int tupan_mlp_infer(int n, const float* feat, const void* packed, float* delta, void* stream)
{
    return mlp_infer(n, feat, packed, delta, static_cast<hipStream_t>(stream));
}

}  // extern "C"
