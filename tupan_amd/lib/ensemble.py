"""Ensembles of independent few-body systems on the GPU.

The particles of all systems live in one device-resident ParticleSystem,
system s occupying rows offsets[s]:offsets[s+1] (at most 64 particles).  An
``Ensemble`` holds the launch plan (tupan_ensemble_create) and runs the
Sakura pair drift or the potential of every system in one launch.
"""
import ctypes

import numpy as np
import torch

from .. import _native
from .hip_backend import current_stream_handle, hip_device


class Ensemble(object):
    def __init__(self, offsets, min_lanes=8):
        offsets = np.ascontiguousarray(offsets, dtype=np.int32)
        if offsets.ndim != 1 or offsets.size < 1 or offsets[0] != 0:
            raise ValueError("offsets must start at 0")
        sizes = np.diff(offsets)
        if np.any(sizes < 0) or np.any(sizes > 64):
            raise ValueError("system sizes must be in [0, 64]")
        hip_device()
        self.lib = _native.lib()
        self.offsets = offsets
        self.nsys = offsets.size - 1
        self.ntot = int(offsets[-1])
        self.min_lanes = int(min_lanes)
        plan = ctypes.c_void_p()
        rc = self.lib.tupan_ensemble_create(
            self.nsys, offsets.ctypes.data_as(ctypes.c_void_p), self.min_lanes,
            ctypes.byref(plan))
        _native.check(rc, "tupan_ensemble_create")
        self._plan = plan
        self._status = torch.zeros(1, dtype=torch.int32, device=hip_device())
        self.counters = None
        self._net = None
        self.fallbacks = None
        ns = ctypes.c_int()
        nt = ctypes.c_int()
        nw = ctypes.c_int()
        _native.check(self.lib.tupan_ensemble_info(
            self._plan, ctypes.byref(ns), ctypes.byref(nt), ctypes.byref(nw)),
            "tupan_ensemble_info")
        self.nwaves = nw.value
        self.system_of = torch.as_tensor(
            np.repeat(np.arange(self.nsys), sizes), device=hip_device())

    def set_guess(self, net=None):
        """Use the Kepler-guess network ``net`` (a KeplerNet, or None for
        the plain guess) in this plan's Sakura calls; its fallbacks are
        counted in ``self.fallbacks`` (a device counter)."""
        self._net = net
        if net is None:
            self.fallbacks = None
            rc = self.lib.tupan_ensemble_set_guess(self._plan, None, None)
        else:
            self.fallbacks = torch.zeros(1, dtype=torch.int64,
                                         device=hip_device())
            rc = self.lib.tupan_ensemble_set_guess(
                self._plan, net.packed.data_ptr(),
                self.fallbacks.data_ptr())
        _native.check(rc, "tupan_ensemble_set_guess")

    def close(self):
        if getattr(self, "_plan", None) is not None and self._plan.value:
            self.lib.tupan_ensemble_destroy(self._plan)
            self._plan = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check_ps(self, ps):
        if ps.n != self.ntot:
            raise ValueError("particle system has %d particles, plan %d"
                             % (ps.n, self.ntot))
        for k in ("mass", "rx", "ry", "rz", "eps2", "vx", "vy", "vz"):
            t = getattr(ps, k)
            if not (t.is_cuda and t.dtype == torch.float64
                    and t.is_contiguous()):
                raise TypeError("ensemble arrays must be contiguous float64 "
                                "device tensors")

    def sakura(self, ps, dt, flag, check=True):
        """dr, dv of every particle from the pair drifts inside its system
        (sakura_kernel.c:5-63); registers drx..dvz on ps.  With
        check=False the solver status is left on the device (no host
        synchronisation; read it with check_status): the form used inside a
        captured graph."""
        self._check_ps(ps)
        for k in ("drx", "dry", "drz", "dvx", "dvy", "dvz"):
            if k not in ps.__dict__:
                ps.register_auxiliary_attribute(k, "real")
        if check:
            self._status.zero_()
        cnt = self.counters.data_ptr() if self.counters is not None else None
        p = [getattr(ps, k).data_ptr() for k in
             ("mass", "rx", "ry", "rz", "eps2", "vx", "vy", "vz")]
        o = [getattr(ps, k).data_ptr() for k in
             ("drx", "dry", "drz", "dvx", "dvy", "dvz")]
        rc = self.lib.tupan_ensemble_sakura(
            self._plan, *p, float(dt), int(flag), *o, cnt,
            self._status.data_ptr(), current_stream_handle())
        _native.check(rc, "tupan_ensemble_sakura")
        if check:
            self.check_status()
        return ps.drx, ps.dry, ps.drz, ps.dvx, ps.dvy, ps.dvz

    def reset_status(self):
        self._status.zero_()

    def check_status(self):
        """Raise if any pair drift since the last reset exhausted the
        solver's bounded step subdivision (one host synchronisation)."""
        if int(self._status.item()) != 0:
            raise _native.TupanHipError(
                "ensemble Sakura: Kepler solver exhausted its bounded step "
                "subdivision for at least one pair")

    def phi(self, ps):
        """Potential of every particle due to its own system."""
        self._check_ps(ps)
        if "phi" not in ps.__dict__:
            ps.register_auxiliary_attribute("phi", "real")
        p = [getattr(ps, k).data_ptr() for k in
             ("mass", "rx", "ry", "rz", "eps2")]
        rc = self.lib.tupan_ensemble_phi(self._plan, *p, ps.phi.data_ptr(),
                                         current_stream_handle())
        _native.check(rc, "tupan_ensemble_phi")
        return ps.phi

    def system_sum(self, values):
        """Per-system sums of a per-particle device array (FP64)."""
        out = torch.zeros(self.nsys, dtype=torch.float64,
                          device=values.device)
        out.index_add_(0, self.system_of, values.to(torch.float64))
        return out

    def energies(self, ps):
        """(kinetic, potential) energy of every system."""
        ke = 0.5 * ps.mass * (ps.vx**2 + ps.vy**2 + ps.vz**2)
        pe = 0.5 * ps.mass * self.phi(ps)
        return self.system_sum(ke), self.system_sum(pe)
