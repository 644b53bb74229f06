"""Learned initial guess for the universal-variable Kepler solver of the
Sakura pair drift.

The solver (universal_kepler_solver.h:360-463 in the reference; kepler.h
here) starts its Laguerre iteration from s0 = dt / r (or, for hyperbolic
orbits with a large step, the Bate-Mueller-White guess).  ``KeplerNet`` is a
4 -> 64 -> 64 -> 1 perceptron with SiLU activations that predicts
delta = s r / dt - 1 for the converged anomaly s from the dimensionless
description of the solve (kepler.h: SolveFeatures):

    tau   = dt sqrt(m / r^3)      step in units of the local dynamical time
    sigma = r.v / sqrt(m r)       radial velocity
    kappa = v^2 r / m             kinetic over potential
    eps   = e2 / (r^2 + e2)       softening share

The network runs in BF16 with FP32 accumulation inside the two-pass Sakura
form (libtupan_hip: tupan_sakura_ex); the iteration then refines the guess to
the solver's usual tolerance (|ds| <= 2^-42) and restarts from the plain
guess if it has not converged after 8 steps, so converged states are the
plain solver's up to that tolerance.

Training runs on the GPU through the library's own kernels (FP32 forward and
backward with deterministic gradient sums, Adam); PyTorch only holds the
device buffers and draws the batches.  The data are the solves the method
itself logs: ``collect`` runs Sakura steps with a solve log attached.
"""
import ctypes

import numpy as np
import torch

from .. import _native
from ..lib.hip_backend import current_stream_handle, hip_device

HIDDEN = 64


def sizes():
    p = ctypes.c_int()
    q = ctypes.c_int()
    b = ctypes.c_int()
    _native.check(_native.lib().tupan_mlp_sizes(ctypes.byref(p),
                                                ctypes.byref(q),
                                                ctypes.byref(b)),
                  "tupan_mlp_sizes")
    return p.value, q.value, b.value


def prepare_inputs(features):
    """Host restatement of the kernels' input transform (log tau, sigma,
    kappa, eps), for the normalisation statistics."""
    f = np.asarray(features, dtype=np.float64)
    x = f.copy()
    x[0] = np.log(np.maximum(f[0], 1e-30))
    return x


class SolveLogBuffer(object):
    """Device log of the near-pair solves of one Sakura call
    (tupan_solve_log): features [4][capacity], target [capacity] and the
    Laguerre steps [capacity] of each pair."""

    def __init__(self, capacity, device=None):
        dev = device if device is not None else hip_device()
        self.capacity = int(capacity)
        self.features = torch.zeros((4, self.capacity), dtype=torch.float32,
                                    device=dev)
        self.target = torch.zeros(self.capacity, dtype=torch.float32,
                                  device=dev)
        self.iterations = torch.zeros(self.capacity, dtype=torch.int32,
                                      device=dev)
        self.count = 0

    def struct(self):
        return _native.SolveLog(self.features.data_ptr(),
                                self.target.data_ptr(),
                                self.iterations.data_ptr(), self.capacity)

    def valid(self):
        """(features [4][k], target [k], iterations [k]) of the logged
        entries whose first solve converged undivided."""
        n = self.count
        t = self.target[:n]
        ok = torch.isfinite(t) & torch.isfinite(self.features[:, :n]).all(0)
        return (self.features[:, :n][:, ok].contiguous(), t[ok].contiguous(),
                self.iterations[:n][ok].contiguous())


class KeplerNet(object):
    """Parameters (FP32, flat, device), normalisation and the packed BF16
    inference block of the guess network."""

    def __init__(self, params=None, norm=None, seed=0, device=None):
        self.device = device if device is not None else hip_device()
        self.nparams, self.nnorm, self.packed_bytes = sizes()
        if params is None:
            params = self.initial_params(seed)
        if norm is None:
            norm = np.array([0, 0, 0, 0, 1, 1, 1, 1, 1, 0], dtype=np.float32)
        self.params = torch.as_tensor(np.asarray(params, dtype=np.float32),
                                      device=self.device).contiguous()
        self.norm = torch.as_tensor(np.asarray(norm, dtype=np.float32),
                                    device=self.device).contiguous()
        if self.params.numel() != self.nparams or \
                self.norm.numel() != self.nnorm:
            raise ValueError("parameter or normalisation size mismatch")
        self.packed = torch.zeros(self.packed_bytes, dtype=torch.uint8,
                                  device=self.device)
        self._m1 = torch.zeros_like(self.params)
        self._m2 = torch.zeros_like(self.params)
        self._grad = torch.zeros_like(self.params)
        self._loss = torch.zeros(1, dtype=torch.float32, device=self.device)
        self.step = 0
        self.pack()

    # -- parameters -------------------------------------------------------
    @staticmethod
    def initial_params(seed):
        """Glorot-uniform weights, zero biases, W1 b1 W2 b2 W3 b3."""
        rng = np.random.RandomState(seed)

        def glorot(fan_out, fan_in):
            lim = np.sqrt(6.0 / (fan_in + fan_out))
            return rng.uniform(-lim, lim, size=(fan_out, fan_in))

        parts = [glorot(HIDDEN, 4).ravel(), np.zeros(HIDDEN),
                 glorot(HIDDEN, HIDDEN).ravel(), np.zeros(HIDDEN),
                 glorot(1, HIDDEN).ravel(), np.zeros(1)]
        return np.concatenate(parts).astype(np.float32)

    def fit_norm(self, features, target):
        """Input standardisation and output scale from training data
        (device or host arrays)."""
        f = features.detach().cpu().numpy() if torch.is_tensor(features) \
            else np.asarray(features)
        t = target.detach().cpu().numpy() if torch.is_tensor(target) \
            else np.asarray(target)
        x = prepare_inputs(f)
        mean = x.mean(axis=1)
        std = x.std(axis=1)
        std[std < 1e-12] = 1.0
        scale = float(np.std(t)) or 1.0
        shift = float(np.mean(t))
        norm = np.concatenate([mean, 1.0 / std, [scale, shift]])
        self.norm.copy_(torch.as_tensor(norm.astype(np.float32)))
        self.pack()

    def pack(self):
        _native.check(_native.lib().tupan_mlp_pack(
            self.params.data_ptr(), self.norm.data_ptr(),
            self.packed.data_ptr(), current_stream_handle()),
            "tupan_mlp_pack")

    # -- evaluation ---------------------------------------------------------
    def _check_features(self, features):
        if not (torch.is_tensor(features) and features.dtype == torch.float32
                and features.is_cuda and features.dim() == 2
                and features.shape[0] == 4 and features.is_contiguous()):
            raise TypeError("features must be a contiguous float32 device "
                            "tensor of shape (4, n)")
        return int(features.shape[1])

    def forward(self, features):
        """delta for each column of features, FP32 arithmetic."""
        n = self._check_features(features)
        out = torch.empty(n, dtype=torch.float32, device=self.device)
        _native.check(_native.lib().tupan_mlp_forward(
            n, features.data_ptr(), self.params.data_ptr(),
            self.norm.data_ptr(), out.data_ptr(), current_stream_handle()),
            "tupan_mlp_forward")
        return out

    def infer(self, features):
        """delta with the packed BF16 network (the Sakura path)."""
        n = self._check_features(features)
        out = torch.empty(n, dtype=torch.float32, device=self.device)
        _native.check(_native.lib().tupan_mlp_infer(
            n, features.data_ptr(), self.packed.data_ptr(), out.data_ptr(),
            current_stream_handle()), "tupan_mlp_infer")
        return out

    def loss_and_grad(self, features, target):
        """(mean squared error of the normalised output, gradient tensor)."""
        n = self._check_features(features)
        if not (target.dtype == torch.float32 and target.numel() == n):
            raise TypeError("target must be float32 with one entry per "
                            "feature column")
        _native.check(_native.lib().tupan_mlp_grad(
            n, features.data_ptr(), target.data_ptr(), self.params.data_ptr(),
            self.norm.data_ptr(), self._grad.data_ptr(),
            self._loss.data_ptr(), current_stream_handle()), "tupan_mlp_grad")
        return self._loss, self._grad

    def adam_step(self, lr, beta1=0.9, beta2=0.999, eps=1e-8):
        self.step += 1
        _native.check(_native.lib().tupan_mlp_adam(
            self.params.data_ptr(), self._grad.data_ptr(),
            self._m1.data_ptr(), self._m2.data_ptr(), lr, beta1, beta2, eps,
            self.step, current_stream_handle()), "tupan_mlp_adam")

    def train(self, features, target, epochs=10, batch=8192, lr=2e-3,
              lr_end=1e-4, seed=0, callback=None):
        """Adam over shuffled mini-batches (the draw is PyTorch's, the
        arithmetic the library's); the learning rate decays geometrically
        from lr to lr_end.  Returns the mean training loss per epoch."""
        n = self._check_features(features)
        gen = torch.Generator(device=self.device)
        gen.manual_seed(seed)
        nb = max(1, n // batch)
        total = epochs * nb
        history = []
        k = 0
        for ep in range(epochs):
            perm = torch.randperm(n, generator=gen, device=self.device)
            acc = torch.zeros(1, dtype=torch.float32, device=self.device)
            for b in range(nb):
                idx = perm[b * batch:(b + 1) * batch]
                fb = features.index_select(1, idx).contiguous()
                tb = target.index_select(0, idx).contiguous()
                loss, _ = self.loss_and_grad(fb, tb)
                acc += loss
                rate = lr * (lr_end / lr)**(k / max(1, total - 1))
                self.adam_step(rate)
                k += 1
            history.append(float(acc) / nb)
            if callback is not None:
                callback(ep, history[-1])
        self.pack()
        return history

    # -- storage ------------------------------------------------------------
    def save(self, path):
        np.savez(path, params=self.params.cpu().numpy(),
                 norm=self.norm.cpu().numpy())

    @classmethod
    def load(cls, path, device=None):
        with np.load(path) as z:
            return cls(params=z["params"], norm=z["norm"], device=device)


def sakura_kernel(prec="float64", **kw):
    """A Sakura extension object on the HIP backend with options."""
    from ..lib import extensions
    return extensions.Sakura("HIP", prec, **kw)


def collect(states, dt=1.0 / 64, flags=(-1, 1), capacity=1 << 22):
    """Log the near-pair solves of Sakura calls on each particle system in
    `states` (one call per flag): returns device (features [4][k],
    target [k], iterations [k]) of the converged first solves."""
    ext = sakura_kernel(path="two-pass")
    log = SolveLogBuffer(capacity)
    ext.kernel.solve_log = log
    feats, targs, iters = [], [], []
    for ps in states:
        for flag in flags:
            ext.calc(ps, ps, dt, flag)
            f, t, it = log.valid()
            feats.append(f.clone())
            targs.append(t.clone())
            iters.append(it.clone())
    return torch.cat(feats, 1), torch.cat(targs), torch.cat(iters)


def collect_trajectory(ps, steps, tau=1.0 / 64, capacity=1 << 22):
    """Advance ps by `steps` Sakura steps (integrator/sakura.py: sakura_step)
    with a solve log on every pair-drift call; returns the logged converged
    first solves as (features [4][k], target [k], iterations [k])."""
    from ..integrator.sakura import sakura_step
    from ..lib import extensions
    ext = sakura_kernel(path="two-pass")
    log = SolveLogBuffer(capacity)
    ext.kernel.solve_log = log
    got = []

    class _Harvest(object):
        def calc(self, ips, jps, dt, flag):
            res = ext.calc(ips, jps, dt, flag)
            got.append(tuple(x.clone() for x in log.valid()))
            return res

    saved = extensions.sakura
    extensions.sakura = _Harvest()
    try:
        for _ in range(steps):
            sakura_step(ps, tau)
    finally:
        extensions.sakura = saved
    return (torch.cat([g[0] for g in got], 1), torch.cat([g[1] for g in got]),
            torch.cat([g[2] for g in got]))
