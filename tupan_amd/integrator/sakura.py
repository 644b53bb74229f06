"""Sakura integrator (tupan/integrator/sakura.py:22-159).

One step: drift tau/2; pairwise Kepler drift with flag -1 over tau/2 and
apply the summed changes; the same with flag +1; drift tau/2
(sakura.py:22-50).  'asakura' picks the shared step from Tstep as
eta / sqrt(1 + max_i(sum_j w2 - max_j w2)) and quantises it to the block
step (sakura.py:100-117).

With device_steps=True the fixed-step 'sakura' method runs its steps through
DeviceSakura: the same operations, with the pair drifts by
tupan_sakura_graph (no host synchronisation), so that Sakura.evolve can
replay a HIP graph of many steps.
"""
import torch

from . import Base
from ..lib import extensions


def sakura_step(ps, tau):
    ps.rx += ps.vx * tau / 2
    ps.ry += ps.vy * tau / 2
    ps.rz += ps.vz * tau / 2

    extensions.sakura.calc(ps, ps, tau / 2, -1)
    ps.rx += ps.drx
    ps.ry += ps.dry
    ps.rz += ps.drz
    ps.vx += ps.dvx
    ps.vy += ps.dvy
    ps.vz += ps.dvz

    extensions.sakura.calc(ps, ps, tau / 2, 1)
    ps.rx += ps.drx
    ps.ry += ps.dry
    ps.rz += ps.drz
    ps.vx += ps.dvx
    ps.vy += ps.dvy
    ps.vz += ps.dvz

    ps.rx += ps.vx * tau / 2
    ps.ry += ps.vy * tau / 2
    ps.rz += ps.vz * tau / 2
    return ps


class DeviceSakura(object):
    """Fixed-step Sakura steps of one particle system with every operation
    on the device: sakura_step's drifts and updates as torch operations on
    the particle arrays, the two pair drifts through tupan_sakura_graph.
    The near-pair list's capacity starts at four times the count of a
    probe call; a step whose list outgrows it is detected from the status
    word after the replay, and the caller restores and retries (evolve)."""

    FIELDS = ("mass", "rx", "ry", "rz", "eps2", "vx", "vy", "vz")
    OUT = ("drx", "dry", "drz", "dvx", "dvy", "dvz")
    STATE = ("rx", "ry", "rz", "vx", "vy", "vz", "time", "nstep", "tstep")

    def __init__(self, ps):
        from .. import _native
        from ..lib.hip_backend import HIPKernel
        k = getattr(extensions.sakura, "kernel", None)
        if not isinstance(k, HIPKernel) or k.prec != "float64":
            raise RuntimeError("device Sakura steps need the HIP backend in "
                               "float64")
        if not ps.mass.is_cuda:
            raise RuntimeError("device Sakura steps need a device-resident "
                               "particle system")
        self.native = _native
        self.lib = _native.lib()
        self.ps = ps
        self.n = ps.n
        for name in self.OUT:
            if name not in ps.__dict__:
                ps.register_auxiliary_attribute(name, "real")
        self.fields = _native.ptr_array([ps.__dict__[f].data_ptr()
                                         for f in self.FIELDS])
        self.outs = _native.ptr_array([ps.__dict__[f].data_ptr()
                                       for f in self.OUT])
        self.status = torch.zeros(1, dtype=torch.int32, device=ps.mass.device)
        self.capacity = 0

    def probe_capacity(self, dt):
        """Near pairs of both pair drifts of a step of dt from the current
        state (tupan_sakura_ex, two-pass form; the outputs are scratch)."""
        import ctypes
        from ..lib.hip_backend import current_stream_handle
        most = 0
        for flag in (-1, 1):
            npairs = ctypes.c_int(-1)
            rc = self.lib.tupan_sakura_ex(
                self.n, self.fields, self.n, self.fields, dt / 2, flag,
                self.outs, 2, None, None, None, ctypes.byref(npairs),
                self.status.data_ptr(), current_stream_handle())
            self.native.check(rc, "tupan_sakura_ex")
            most = max(most, npairs.value)
        self.check()
        self.capacity = max(4 * most, 4096)

    def pair_drift(self, dt, flag):
        from ..lib.hip_backend import current_stream_handle
        rc = self.lib.tupan_sakura_graph(
            self.n, self.fields, self.n, self.fields, dt, flag, self.outs,
            self.capacity, None, self.status.data_ptr(),
            current_stream_handle())
        self.native.check(rc, "tupan_sakura_graph")

    def step(self, tau):
        """sakura_step and Sakura.do_step's bookkeeping, on the device."""
        ps = self.ps
        ps.rx += ps.vx * tau / 2
        ps.ry += ps.vy * tau / 2
        ps.rz += ps.vz * tau / 2
        for flag in (-1, 1):
            self.pair_drift(tau / 2, flag)
            ps.rx += ps.drx
            ps.ry += ps.dry
            ps.rz += ps.drz
            ps.vx += ps.dvx
            ps.vy += ps.dvy
            ps.vz += ps.dvz
        ps.rx += ps.vx * tau / 2
        ps.ry += ps.vy * tau / 2
        ps.rz += ps.vz * tau / 2
        ps.tstep[...] = tau
        ps.time += tau
        ps.nstep += 1

    def check(self):
        """0, or 2 when a near-pair list outgrew its capacity (the state is
        then invalid); raises on a Kepler failure (status bit 1)."""
        st = int(self.status.item())
        if st & 1:
            raise self.native.TupanHipError(
                "tupan_sakura_graph: a pair exhausted the Kepler solver's "
                "bounded step subdivision")
        return st & 2

    def snapshot(self):
        return {k: self.ps.__dict__[k].clone() for k in self.STATE}

    def restore(self, snap):
        for k, v in snap.items():
            self.ps.__dict__[k].copy_(v)

    def capture(self, tau, nsteps):
        graph = torch.cuda.CUDAGraph()
        side = torch.cuda.Stream()
        side.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(side):
            with torch.cuda.graph(graph, stream=side):
                for _ in range(nsteps):
                    self.step(tau)
        torch.cuda.current_stream().wait_stream(side)
        return graph


class Sakura(Base):
    PROVIDED_METHODS = ["sakura", "asakura"]

    def __init__(self, eta, time, ps, method, device_steps=False, **kwargs):
        super(Sakura, self).__init__(eta, time, ps, **kwargs)
        self.method = method
        self.device_steps = device_steps
        if device_steps and "asakura" in method:
            raise NotImplementedError(
                "device Sakura steps are built for the fixed-step 'sakura' "
                "method")
        self._dev = None

    def _device(self):
        if self._dev is None or self._dev.ps is not self.ps:
            self._dev = DeviceSakura(self.ps)
        return self._dev

    def evolve_step(self, t_end):
        """integrator/__init__.py:80-99; with device_steps one device step."""
        if not self.device_steps:
            return super(Sakura, self).evolve_step(t_end)
        if not self.is_initialized:
            self.initialize(t_end)
        tau = self.get_base_tstep(t_end)
        dev = self._device()
        if dev.capacity == 0:
            dev.probe_capacity(tau)
        self._eager_step(dev, tau)
        if self.reporter:
            self.reporter.diagnostic_report(self.ps)

    def _eager_step(self, dev, tau):
        while True:
            snap = dev.snapshot()
            dev.status.zero_()
            dev.step(tau)
            if not dev.check():
                break
            dev.restore(snap)
            dev.capacity *= 2
        type(self.ps).t_curr += tau

    def evolve(self, t_end, steps_per_graph=16):
        """Advance to t_end by device steps of eta, replaying a HIP graph of
        steps_per_graph steps while a whole graph fits before t_end and
        finishing with single steps; a replay whose near-pair list outgrew
        its capacity is undone and run again with twice the capacity.
        Returns the number of replays."""
        if not self.device_steps:
            raise RuntimeError("evolve needs device_steps=True")
        if steps_per_graph < 1:
            raise ValueError("steps_per_graph must be at least 1")
        if not self.is_initialized:
            self.initialize(t_end)
        dev = self._device()
        tau = self.eta
        ps = self.ps
        replays = 0
        graph, graph_cap = None, -1
        if dev.capacity == 0:
            dev.probe_capacity(tau)
        while abs(t_end) - abs(ps.t_curr) >= steps_per_graph * abs(tau):
            if graph is None or graph_cap != dev.capacity:
                self._eager_step(dev, tau)   # sizes workspaces for capture
                graph, graph_cap = dev.capture(tau, steps_per_graph), dev.capacity
                continue
            snap = dev.snapshot()
            dev.status.zero_()
            graph.replay()
            if dev.check():
                dev.restore(snap)
                dev.capacity *= 2
                continue
            for _ in range(steps_per_graph):
                type(ps).t_curr += tau
            replays += 1
        while abs(ps.t_curr) < abs(t_end):
            self._eager_step(dev, self.get_base_tstep(t_end))
        if self.reporter:
            self.reporter.diagnostic_report(ps)
        return replays

    def get_sakura_tstep(self, ps, eta, tau):
        ps.set_tstep(ps, eta)
        iw2_a = (eta / ps.tstep)**2
        iw2_b = (eta / ps.tstepij)**2
        diw2 = iw2_a - iw2_b
        w2_sakura = float(diw2.max())
        dt_sakura = eta / (1 + w2_sakura)**0.5
        ps.tstep[...] = dt_sakura
        return self.get_min_block_tstep(ps, tau)

    def do_step(self, ps, tau):
        if "asakura" in self.method:
            tau = self.get_sakura_tstep(ps, self.eta, tau)
        ps = sakura_step(ps, tau)
        type(ps).t_curr += tau
        ps.tstep[...] = tau
        ps.time += tau
        ps.nstep += 1
        return ps
