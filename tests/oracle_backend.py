"""Test-only backend: the reference kernel-object protocol on the CPU oracle.

Lets the CPU test suite run the host logic (extension layer, particle
system, integrators, initial conditions) without a GPU, on particle systems
whose arrays live on the CPU.  It is installed only by tests; the product
ships the HIP backend alone.
"""
from collections import namedtuple

import numpy as np
import torch

from oracle import SPECS, Oracle

_NAMES = {"phi_kernel": "phi", "acc_kernel": "acc",
          "acc_jerk_kernel": "acc_jerk", "snap_crackle_kernel": "snap_crackle",
          "tstep_kernel": "tstep", "sakura_kernel": "sakura",
          "kepler_solver_kernel": "kepler", "nreg_Xkernel": "nreg_x",
          "nreg_Vkernel": "nreg_v", "pnacc_kernel": "pnacc"}
_FIELDS = ("mass", "rx", "ry", "rz", "eps2", "vx", "vy", "vz")


def _np(x):
    if isinstance(x, torch.Tensor):
        return x.detach().cpu().numpy()
    return np.asarray(x)


class OracleKernel(object):
    calls = []  # (name, ni, nj) log, for tests of call patterns

    def __init__(self, prec, name):
        self.name = _NAMES[name]
        self.prec = prec
        self.orc = Oracle(prec)
        ident = lambda x: x  # noqa: E731
        T = namedtuple("Types", ["c_int", "c_int_p", "c_uint", "c_uint_p",
                                 "c_real", "c_real_p"])
        self.cty = T(ident, ident, int, ident, float, ident)

    def set_gsize(self, ni, nj):
        pass

    def set_args(self, args, start=0):
        self.args = args

    def map_buffers(self, arrays, buffers):
        return arrays

    def run(self):
        a = self.args
        if self.name == "kepler":
            ps = dict(zip(_FIELDS[:8], [_np(x) for x in a[:8]]))
            res = self.orc.kepler(ps, a[8])
            for out, r in zip(a[9:15], res):
                if isinstance(out, torch.Tensor):
                    out[:2] = torch.as_tensor(r[:2], dtype=out.dtype)
                else:
                    out[:2] = r[:2]
            OracleKernel.calls.append((self.name, 2, 2))
            return
        ifl, jfl, sfl, ofl = SPECS[self.name]
        k = 0
        ni = int(a[k]); k += 1
        ips = dict(zip(ifl, [_np(x)[:ni] for x in a[k:k + len(ifl)]]))
        k += len(ifl)
        nj = int(a[k]); k += 1
        jps = dict(zip(jfl, [_np(x)[:nj] for x in a[k:k + len(jfl)]]))
        k += len(jfl)
        sc = a[k:k + len(sfl)]
        k += len(sfl)
        outs = a[k:k + len(ofl)]
        res = self.orc.pairwise(self.name, ips, jps, *sc)
        for out, r in zip(outs, res):
            if isinstance(out, torch.Tensor):
                out[:ni] = torch.as_tensor(r, dtype=out.dtype)
            else:
                out[:ni] = r
        OracleKernel.calls.append((self.name, ni, nj))


def install_oracle(prec="float64"):
    from tupan_amd.lib import extensions
    extensions.install(OracleKernel, prec)


def install_hip(prec="float64", **kw):
    from tupan_amd.lib import extensions
    extensions.install("HIP", prec, **kw)
