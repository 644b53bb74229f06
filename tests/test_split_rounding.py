"""Why the split-FP32 kernels form R^2 = fma(dz, dz, fma(dy, dy, fma(dx, dx,
e2))) (CPU, numpy float32 with a correctly rounded 1/sqrt).

On a Plummer sphere with eps = 1e-4 (pair e2 = 1e-8), rounding r^2 to FP32
and then adding e2 drops e2 for every pair whose e2 is below half an ulp of
r^2, so the far forces come out too strong by ~1.5 e2 / R^2 on average: the
virial W = sum m_i r_i . a_i (the potential energy) is off by ~+1.2e-8, which
an integration turns into |dE/E| ~ 1e-8 |dU/E|.  Entering e2 before the
first rounding leaves every rounding unbiased.  The GPU test
test_gpu_precision.py::test_split32_force_law_unbiased checks the kernels
themselves.
"""
import numpy as np

f32, f64 = np.float32, np.float64


def plummer(n, seed):
    rng = np.random.default_rng(seed)
    a = 3 * np.pi / 16
    x = rng.uniform(0.0, 0.999, n)
    r = a / np.sqrt(x ** (-2.0 / 3.0) - 1.0)
    mu = rng.uniform(-1, 1, n)
    ph = rng.uniform(0, 2 * np.pi, n)
    s = np.sqrt(1 - mu * mu)
    return np.stack([r * s * np.cos(ph), r * s * np.sin(ph), r * mu])


def virial_error(order, n=8192, ni=1024, eps=1e-4, seed=3):
    x = plummer(n, seed)
    m = np.full(n, 1.0 / n)
    e2 = f32(eps * eps / 2) + f32(eps * eps / 2)          # pair e2, as in FP32
    rows = np.random.default_rng(seed + 1).choice(n, ni, replace=False)
    w32 = w64 = 0.0
    for i in rows:
        d = x[:, [i]] - x                                 # exact in FP64
        dx = d.astype(f32)                                # one rounding each
        if order == "e2_first":
            R2 = dx[0].astype(f64) * dx[0] + e2           # fma(dx, dx, e2)
            R2 = R2.astype(f32).astype(f64)
        else:                                             # r^2, then + e2
            R2 = (dx[0].astype(f64) * dx[0]).astype(f32).astype(f64)
        for k in (1, 2):                                  # fma chain
            R2 = (dx[k].astype(f64) * dx[k] + R2).astype(f32).astype(f64)
        if order == "e2_first":
            mask = R2 > e2
        else:
            mask = R2 > 0
            R2 = (R2 + e2).astype(f32).astype(f64)
        y = np.where(mask, 1.0 / np.sqrt(np.where(mask, R2, 1.0)), 0.0)
        y = y.astype(f32).astype(f64)                     # correctly rounded
        q3 = (y * y).astype(f32).astype(f64) * (y * m).astype(f32)
        q3 = q3.astype(f32).astype(f64)
        a32 = -(q3 * dx.astype(f64)).sum(1)
        r2 = (d * d).sum(0)
        yy = np.where(r2 > 0, 1.0 / np.sqrt(r2 + f64(e2)), 0.0)
        a64 = -(m * yy ** 3 * d).sum(1)
        w32 += m[i] * (x[:, i] @ a32)
        w64 += m[i] * (x[:, i] @ a64)
    return w32 / w64 - 1


def test_rounding_r2_before_e2_biases_the_force_law():
    beta = virial_error("r2_first")
    print("r2 then + e2: virial scale error %.3e" % beta)
    assert 6e-9 <= beta <= 2e-8


def test_e2_first_is_unbiased():
    beta = virial_error("e2_first")
    print("e2 first: virial scale error %.3e" % beta)
    assert abs(beta) <= 2e-9
