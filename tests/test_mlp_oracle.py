"""The CPU oracle of the Kepler-guess network (oracle/mlp.py), pinned by
what mathematics fixes: analytic gradients against central finite
differences, Adam's first step in closed form, BF16 rounding on known
values, and the BF16 emulation staying within BF16 resolution of the exact
forward pass."""
import numpy as np

from oracle import mlp


def random_case(n=40, seed=0):
    rng = np.random.RandomState(seed)
    params = rng.normal(0, 0.3, 4545)
    feats = np.stack([np.exp(rng.uniform(-6, 1, n)), rng.normal(0, 1, n),
                      rng.uniform(0, 4, n), rng.uniform(0, 0.3, n)])
    norm = np.array([-2.5, 0.1, 1.5, 0.1, 0.6, 1.1, 0.8, 3.0, 0.02, -0.01])
    target = rng.normal(-0.01, 0.02, n)
    return params, norm, feats, target


def test_gradient_matches_finite_differences():
    params, norm, feats, target = random_case()
    loss, grad = mlp.loss_and_grad(params, norm, feats, target)
    assert grad.shape == params.shape
    rng = np.random.RandomState(1)
    # every block: W1 b1 W2 b2 W3 b3
    picks = np.concatenate([rng.choice(256, 8, replace=False),
                            256 + rng.choice(64, 6, replace=False),
                            320 + rng.choice(4096, 12, replace=False),
                            4416 + rng.choice(64, 6, replace=False),
                            4480 + rng.choice(64, 6, replace=False), [4544]])
    h = 1e-6
    for k in picks:
        p1 = params.copy()
        p1[k] += h
        p0 = params.copy()
        p0[k] -= h
        fd = (mlp.loss_and_grad(p1, norm, feats, target)[0]
              - mlp.loss_and_grad(p0, norm, feats, target)[0]) / (2 * h)
        assert abs(fd - grad[k]) <= 1e-6 * max(1e-3, abs(grad[k])), \
            (k, fd, grad[k])
    # the loss is the mean squared error of the normalised output
    y = mlp.forward(params, norm, feats)
    t = (target - norm[9]) / norm[8]
    assert abs(loss - np.mean((y - t)**2)) <= 1e-12 * loss


def test_adam_first_step_closed_form():
    rng = np.random.RandomState(2)
    p = rng.normal(size=50)
    g = rng.normal(size=50)
    z = np.zeros(50)
    p1, m1, m2 = mlp.adam(p, g, z, z, 1e-3, 0.9, 0.999, 1e-8, 1)
    np.testing.assert_allclose(p1, p - 1e-3 * g / (np.abs(g) + 1e-8),
                               rtol=1e-12, atol=1e-15)
    np.testing.assert_allclose(m1, 0.1 * g)
    np.testing.assert_allclose(m2, 0.001 * g * g)


def test_bf16_rounding():
    vals = np.array([1.0, 1 + 2.0**-8, 1 + 3 * 2.0**-9, -2.5, 3.0e-3,
                     1 + 2.0**-7 + 2.0**-8], dtype=np.float32)
    # 3e-3 = 1.536 * 2^-9 -> 197/128 * 2^-9 (7 stored mantissa bits)
    want = [1.0, 1.0, 1 + 2.0**-7, -2.5, 197 / 128 * 2.0**-9, 1 + 2.0**-6]
    np.testing.assert_array_equal(mlp.bf16(vals), want)


def test_bf16_emulation_within_resolution():
    params, norm, feats, target = random_case(200, 3)
    params = params * 0.5
    exact = mlp.delta(params, norm, feats)
    emu = mlp.delta_bf16(params, norm, feats)
    scale = np.max(np.abs(exact - norm[9])) + 1e-12
    err = np.max(np.abs(emu - exact)) / scale
    assert 1e-5 < err < 0.1, err
