"""CPU oracle of the Kepler-guess network -- TEST INFRASTRUCTURE ONLY.

Float64 numpy restatement of the network of tupan_amd/csrc/mlp.hip (the
reference has no learned part; this pins the GPU kernels): forward pass,
the mean-squared-error loss and its analytic gradient, one Adam step, and an
emulation of the BF16 inference arithmetic (BF16 operands, wide
accumulation).  Layout: params = W1[64][4] b1[64] W2[64][64] b2[64] W3[64]
b3; norm = mean[4], inv_std[4], out_scale, out_shift; inputs
x = ((log tau, sigma, kappa, eps) - mean) * inv_std.
"""
import numpy as np

H = 64
SHAPES = (("W1", (H, 4)), ("b1", (H,)), ("W2", (H, H)), ("b2", (H,)),
          ("W3", (H,)), ("b3", (1,)))


def unpack(params):
    p = np.asarray(params, dtype=np.float64)
    out = {}
    k = 0
    for name, shape in SHAPES:
        size = int(np.prod(shape))
        out[name] = p[k:k + size].reshape(shape)
        k += size
    assert k == p.size
    return out


def inputs(features, norm):
    f = np.asarray(features, dtype=np.float64)
    x = f.copy()
    x[0] = np.log(np.maximum(f[0], 1e-30))
    norm = np.asarray(norm, dtype=np.float64)
    return (x - norm[:4, None]) * norm[4:8, None]


def silu(z):
    return z / (1.0 + np.exp(-z))


def dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def forward(params, norm, features, full=False):
    """Network output y (normalised) for features [4][n]; with full=True
    also the intermediate values."""
    P = unpack(params)
    x = inputs(features, norm)
    z1 = P["W1"] @ x + P["b1"][:, None]
    h1 = silu(z1)
    z2 = P["W2"] @ h1 + P["b2"][:, None]
    h2 = silu(z2)
    y = P["W3"] @ h2 + P["b3"][0]
    if full:
        return y, (x, z1, h1, z2, h2)
    return y


def delta(params, norm, features):
    norm = np.asarray(norm, dtype=np.float64)
    return forward(params, norm, features) * norm[8] + norm[9]


def loss_and_grad(params, norm, features, target):
    """Mean squared error of y against (target - shift) / scale and its
    gradient with respect to the flat parameters."""
    norm = np.asarray(norm, dtype=np.float64)
    P = unpack(params)
    y, (x, z1, h1, z2, h2) = forward(params, norm, features, full=True)
    t = (np.asarray(target, dtype=np.float64) - norm[9]) / norm[8]
    n = y.size
    r = y - t
    loss = float(np.mean(r * r))
    dy = 2.0 * r / n
    gW3 = h2 @ dy
    gb3 = np.array([dy.sum()])
    dz2 = (P["W3"][:, None] * dy[None, :]) * dsilu(z2)
    gW2 = dz2 @ h1.T
    gb2 = dz2.sum(1)
    dz1 = (P["W2"].T @ dz2) * dsilu(z1)
    gW1 = dz1 @ x.T
    gb1 = dz1.sum(1)
    grad = np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gW3, gb3])
    return loss, grad


def adam(params, grad, m1, m2, lr, beta1, beta2, eps, step):
    m1 = beta1 * m1 + (1 - beta1) * grad
    m2 = beta2 * m2 + (1 - beta2) * grad * grad
    c1 = 1 - beta1**step
    c2 = 1 - beta2**step
    return params - lr * (m1 / c1) / (np.sqrt(m2 / c2) + eps), m1, m2


def bf16(a):
    """Round float32 values to the nearest BF16 (ties to even)."""
    u = np.asarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16 << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def delta_bf16(params, norm, features):
    """The BF16 inference arithmetic: inputs, weights and hidden
    activations rounded to BF16, sums and biases in wide precision."""
    norm = np.asarray(norm, dtype=np.float64)
    P = unpack(np.asarray(params, dtype=np.float32).astype(np.float64))
    x = bf16(inputs(features, norm).astype(np.float32))
    z1 = bf16(P["W1"]) @ x + np.float32(P["b1"])[:, None]
    h1 = bf16(silu(z1))
    z2 = bf16(P["W2"]) @ h1 + np.float32(P["b2"])[:, None]
    h2 = bf16(silu(z2))
    y = bf16(P["W3"]) @ h2 + P["b3"][0]
    return y * norm[8] + norm[9]
