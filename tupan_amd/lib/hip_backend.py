"""HIP backend: the reference's kernel-object protocol on libtupan_hip.so.

The reference backends (tupan/lib/cffi_backend.py:90-126 and
tupan/lib/opencl_backend.py:104-171) expose, per kernel:

  cty          converters c_int, c_int_p, c_uint, c_uint_p, c_real, c_real_p
  set_gsize    (ni, nj) launch geometry
  set_args     the converted argument list
  run          execute synchronously
  map_buffers  (arrays, buffers) -> arrays, after the run

``HIPKernel`` implements the same protocol.  ``c_real_p`` accepts

  * a torch tensor on a HIP device: used in place (zero copy), so particle
    arrays stay resident in HBM across calls;
  * a numpy array (or CPU tensor): staged to the device for the call and,
    when it is an output, copied back by ``map_buffers``, like the OpenCL
    backend's per-call buffers (opencl_backend.py:115-130,149-164).

Launches are enqueued on torch's current stream.  There is no CPU path: a
missing library or device raises.
"""
import ctypes

import numpy as np

try:
    import torch
except ImportError:  # pragma: no cover - torch is part of the platform
    torch = None

from .. import _native
from .utils import ctype

# name -> (number of i fields, number of j fields, scalar kinds, outputs)
KERNEL_SPECS = {
    "phi_kernel": (5, 5, (), 1),
    "acc_kernel": (5, 5, (), 3),
    "acc_jerk_kernel": (8, 8, (), 6),
    "snap_crackle_kernel": (14, 14, (), 6),
    "tstep_kernel": (8, 8, ("real",), 2),
    "sakura_kernel": (8, 8, ("real", "int"), 6),
    "kepler_solver_kernel": (8, 0, ("real",), 6),
    "nreg_Xkernel": (8, 8, ("real",), 7),
    "nreg_Vkernel": (7, 7, ("real",), 4),
    "pnacc_kernel": (8, 8, ("uint",) + ("real",) * 7, 3),
}
PAIRWISE_IDS = {
    "phi_kernel": _native.K_PHI,
    "acc_kernel": _native.K_ACC,
    "acc_jerk_kernel": _native.K_ACC_JERK,
    "snap_crackle_kernel": _native.K_SNAP_CRACKLE,
    "tstep_kernel": _native.K_TSTEP,
    "nreg_Xkernel": _native.K_NREG_X,
    "nreg_Vkernel": _native.K_NREG_V,
    "pnacc_kernel": _native.K_PNACC,
}


def _require_torch():
    if torch is None:
        raise _native.TupanHipError("torch (ROCm build) is required")


def hip_device():
    _require_torch()
    if not torch.cuda.is_available():
        raise _native.TupanHipError(
            "no HIP device is visible; tupan_amd kernels run on the GPU only")
    return torch.device("cuda", torch.cuda.current_device())


def current_stream_handle():
    return torch.cuda.current_stream().cuda_stream


class DeviceBuffer(object):
    """A device pointer for one argument, with its staging copy if any."""
    __slots__ = ("ptr", "tensor", "host", "n")

    def __init__(self, ptr, tensor, host, n):
        self.ptr = ptr
        self.tensor = tensor
        self.host = host
        self.n = n


class HIPKernel(object):
    """One extension kernel on libtupan_hip.so (reference protocol)."""

    def __init__(self, prec, name, mode=None, path=None, guess=None):
        if name not in KERNEL_SPECS:
            raise ValueError("unknown kernel %r" % (name,))
        if prec not in ("float32", "float64"):
            raise KeyError(prec)
        self.lib = _native.lib()
        self.name = name
        self.prec = prec
        self.mode = ctype.force_mode if mode is None else mode
        if self.mode not in _native.MODES:
            raise ValueError("unknown force mode %r" % (self.mode,))
        self.np_dtype = np.float64 if prec == "float64" else np.float32
        self._cache = {}
        self._status = None
        self.args = None
        self.ni = self.nj = 0
        # Sakura options: optional device uint64 counters (4, or 5 with the
        # guess fallbacks), the form (None automatic, "one-pass",
        # "two-pass"), a Kepler-guess network (tupan_amd.learned.KeplerNet)
        # and a solve log (tupan_amd.learned.SolveLogBuffer)
        self.counters = None
        if path not in (None, "one-pass", "two-pass"):
            raise ValueError("unknown Sakura path %r" % (path,))
        self.path = path
        self.guess = guess
        self.solve_log = None
        self.near_pairs = None
        conv = self._convert
        from collections import namedtuple
        types = namedtuple("Types", ["c_int", "c_int_p", "c_uint", "c_uint_p",
                                     "c_real", "c_real_p"])
        self.cty = types(c_int=int, c_int_p=conv, c_uint=int, c_uint_p=conv,
                         c_real=float, c_real_p=conv)

    # -- protocol -----------------------------------------------------------
    def set_gsize(self, ni, nj):
        self.ni, self.nj = int(ni), int(nj)
        self._cache = {}

    def allocate_local_memory(self, numbufs, sctype):
        return []

    def set_args(self, args, start=0):
        self.args = list(args)

    def run(self):
        spec = KERNEL_SPECS[self.name]
        if self.name == "kepler_solver_kernel":
            self._run_kepler()
        elif self.name == "sakura_kernel":
            self._run_sakura(spec)
        else:
            self._run_pairwise(spec)
        self._cache = {}

    def map_buffers(self, arrays, buffers):
        staged = [(a, b) for a, b in zip(arrays, buffers)
                  if b.host is not None]
        if staged:
            torch.cuda.current_stream().synchronize()
            for a, b in staged:
                out = b.tensor.detach().cpu().numpy()
                if isinstance(a, np.ndarray):
                    a[...] = out.astype(a.dtype, copy=False)
                else:
                    a.copy_(torch.from_numpy(out))
        return arrays

    # -- argument conversion -----------------------------------------------
    def _convert(self, x):
        key = id(x)
        hit = self._cache.get(key)
        if hit is not None and hit[0] is x:
            return hit[1]
        _require_torch()
        want = torch.float64 if self.prec == "float64" else torch.float32
        if isinstance(x, torch.Tensor) and x.is_cuda:
            if x.dtype != want:
                raise TypeError("%s: expected %s device array, got %s"
                                % (self.name, want, x.dtype))
            if not x.is_contiguous():
                raise ValueError("%s: device arrays must be contiguous"
                                 % self.name)
            buf = DeviceBuffer(x.data_ptr(), x, None, x.numel())
        else:
            dev = hip_device()
            host = x.numpy() if isinstance(x, torch.Tensor) else np.asarray(x)
            arr = np.ascontiguousarray(host, dtype=self.np_dtype)
            t = torch.from_numpy(arr).to(dev)
            buf = DeviceBuffer(t.data_ptr(), t, x, t.numel())
        self._cache[key] = (x, buf)
        return buf

    def _split_args(self, spec):
        nin_i, nin_j, scal, nout = spec
        a = self.args
        k = 0
        ni = int(a[k]); k += 1
        ib = a[k:k + nin_i]; k += nin_i
        nj = int(a[k]); k += 1
        jb = a[k:k + nin_j]; k += nin_j
        sc = a[k:k + len(scal)]; k += len(scal)
        ob = a[k:k + nout]
        for b in list(ib) + list(ob):
            if b.n < ni:
                raise ValueError("%s: i/output arrays shorter than ni"
                                 % self.name)
        for b in jb:
            if b.n < nj:
                raise ValueError("%s: j arrays shorter than nj" % self.name)
        return ni, ib, nj, jb, sc, ob

    # -- launches -------------------------------------------------------------
    def _run_pairwise(self, spec):
        ni, ib, nj, jb, sc, ob = self._split_args(spec)
        if ni == 0:
            return
        hip_device()
        eta = float(sc[0]) if sc else 0.0
        ip = _native.ptr_array([b.ptr for b in ib])
        jp = _native.ptr_array([b.ptr for b in jb])
        op = _native.ptr_array([b.ptr for b in ob])
        kid = PAIRWISE_IDS[self.name]
        stream = current_stream_handle()
        if kid == _native.K_PNACC:
            # (order, 1/c, ..., 1/c^7) as the reference passes them
            # (extensions.py:424-430)
            order = int(sc[0])
            if order < 0:
                raise ValueError("pnacc_kernel: negative PN order")
            if self.prec == "float64":
                inv = (ctypes.c_double * 7)(*[float(x) for x in sc[1:8]])
                rc = self.lib.tupan_pnacc(_native.MODES[self.mode], ni, ip, nj,
                                          jp, order, inv, op, stream)
            else:
                inv = (ctypes.c_float * 7)(*[float(x) for x in sc[1:8]])
                rc = self.lib.tupan_pnacc_f32(0, ni, ip, nj, jp, order, inv,
                                              op, stream)
            _native.check(rc, self.name)
            return
        if self.prec == "float64":
            rc = self.lib.tupan_pairwise(kid, _native.MODES[self.mode], ni, ip,
                                         nj, jp, op, eta, stream)
        else:
            rc = self.lib.tupan_pairwise_f32(kid, 0, ni, ip, nj, jp, op, eta,
                                             stream)
        _native.check(rc, self.name)

    def _status_buffer(self):
        if self._status is None:
            self._status = torch.zeros(1, dtype=torch.int32,
                                       device=hip_device())
        self._status.zero_()
        return self._status

    def _as64(self, bufs):
        if self.prec == "float64":
            return [b.tensor for b in bufs]
        return [b.tensor.double() for b in bufs]

    def _run_sakura(self, spec):
        ni, ib, nj, jb, sc, ob = self._split_args(spec)
        if ni == 0:
            return
        hip_device()
        dt, flag = float(sc[0]), int(sc[1])
        it = self._as64(ib)
        jt = self._as64(jb)
        if self.prec == "float64":
            ot = [b.tensor for b in ob]
        else:
            ot = [torch.empty(ni, dtype=torch.float64, device=b.tensor.device)
                  for b in ob]
        status = self._status_buffer()
        cnt = self.counters.data_ptr() if self.counters is not None else None
        if self.path is None and self.guess is None and self.solve_log is None:
            rc = self.lib.tupan_sakura(
                ni, *[t.data_ptr() for t in it], nj,
                *[t.data_ptr() for t in jt], dt, flag,
                *[t.data_ptr() for t in ot], cnt, status.data_ptr(),
                current_stream_handle())
        else:
            if self.counters is not None and self.counters.numel() < 5:
                raise ValueError("Sakura options need 5 counters")
            path = {None: 0, "one-pass": 1, "two-pass": 2}[self.path]
            net = self.guess.packed.data_ptr() if self.guess is not None \
                else None
            log = self.solve_log.struct() if self.solve_log is not None \
                else None
            near = ctypes.c_int(-1)
            rc = self.lib.tupan_sakura_ex(
                ni, _native.ptr_array([t.data_ptr() for t in it]), nj,
                _native.ptr_array([t.data_ptr() for t in jt]), dt, flag,
                _native.ptr_array([t.data_ptr() for t in ot]), path, net,
                ctypes.byref(log) if log is not None else None, cnt,
                ctypes.byref(near), status.data_ptr(),
                current_stream_handle())
            self.near_pairs = near.value
            if self.solve_log is not None:
                self.solve_log.count = max(0, min(near.value,
                                                  self.solve_log.capacity))
        _native.check(rc, self.name)
        if int(status.item()) != 0:
            raise _native.TupanHipError(
                "sakura_kernel: Kepler solver exhausted its bounded step "
                "subdivision for at least one pair")
        if self.prec != "float64":
            for b, t in zip(ob, ot):
                b.tensor.copy_(t)

    def _run_kepler(self):
        a = self.args
        ib = a[0:8]
        dt = float(a[8])
        ob = a[9:15]
        if any(b.n < 2 for b in ib):
            raise ValueError("kepler_solver_kernel needs two particles")
        hip_device()
        dev = ib[0].tensor.device
        t = self._as64(ib)
        m, rx, ry, rz, e2, vx, vy, vz = t
        state = [rx, ry, rz, vx, vy, vz]
        aliased = all(b.tensor.data_ptr() == s.data_ptr()
                      for b, s in zip(ob, state))
        if not aliased:
            # the kernel updates in place: work on copies of the state
            rx, ry, rz, vx, vy, vz = [s[:2].clone() for s in state]
            m, e2 = m[:2], e2[:2]
        idx = torch.tensor([0, 1], dtype=torch.int32, device=dev)
        status = self._status_buffer()
        rc = self.lib.tupan_kepler_pairs(
            1, idx.data_ptr(), m.data_ptr(), rx.data_ptr(), ry.data_ptr(),
            rz.data_ptr(), e2.data_ptr(), vx.data_ptr(), vy.data_ptr(),
            vz.data_ptr(), dt, status.data_ptr(), current_stream_handle())
        _native.check(rc, self.name)
        if int(status.item()) != 0:
            raise _native.TupanHipError(
                "kepler_solver_kernel: Kepler solver exhausted its bounded "
                "step subdivision")
        # the reference extension layer passes the inputs as outputs
        # (extensions.py:642-646); otherwise copy the new state out
        if not aliased:
            for b, r in zip(ob, (rx, ry, rz, vx, vy, vz)):
                b.tensor[:2].copy_(r[:2])
