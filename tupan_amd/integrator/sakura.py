"""Sakura integrator (tupan/integrator/sakura.py:22-159).

One step: drift tau/2; pairwise Kepler drift with flag -1 over tau/2 and
apply the summed changes; the same with flag +1; drift tau/2
(sakura.py:22-50).  'asakura' picks the shared step from Tstep as
eta / sqrt(1 + max_i(sum_j w2 - max_j w2)) and quantises it to the block
step (sakura.py:100-117).

With device_steps=True every step is one call of tupan_sakura_step (the
same operations on the device, the step size computed there from a control
block, the pair drifts in their graph-safe two-pass form), so that
Sakura.evolve can replay a HIP graph of many steps.
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
    """Sakura steps of one integrator run by tupan_sakura_step: the particle
    arrays stay in place, a control block (t_curr, t_end, eta, tau, status,
    dt) and, for 'asakura', a 2 n scratch array live on the device.  The
    pair drifts' near-pair list has a fixed capacity, four times the count
    of a probe call to start with; a step whose list outgrows it sets bit 2
    of the status word, and the caller restores the state saved before the
    step (or replay) and runs it again with twice the capacity."""

    ROLES = ("mass", "eps2", "rx", "ry", "rz", "vx", "vy", "vz",
             "drx", "dry", "drz", "dvx", "dvy", "dvz",
             "tstep", "tstepij", "time", "nstep")
    OUT = ("drx", "dry", "drz", "dvx", "dvy", "dvz")
    FIELDS = ("mass", "rx", "ry", "rz", "eps2", "vx", "vy", "vz")
    STATE = ("rx", "ry", "rz", "vx", "vy", "vz", "time", "nstep", "tstep",
             "tstepij")
    STATUS = {1: "non-finite time, t_end or eta",
              2: "Tstep gave no finite maximum of w2_a - w2_b",
              3: "no power-of-two block step aligned with the time"}

    def __init__(self, integ):
        from .. import _native
        from ..lib.hip_backend import HIPKernel
        ps = integ.ps
        k = getattr(extensions.sakura, "kernel", None)
        kt = getattr(extensions.tstep, "kernel", None)
        if (not isinstance(k, HIPKernel) or k.prec != "float64"
                or not isinstance(kt, HIPKernel)):
            raise RuntimeError("device Sakura steps need the HIP backend in "
                               "float64")
        if not ps.mass.is_cuda:
            raise RuntimeError("device Sakura steps need a device-resident "
                               "particle system")
        self.native = _native
        self.lib = _native.lib()
        self.integ = integ
        self.mode = _native.MODES[kt.mode]
        self.adaptive = 1 if "asakura" in integ.method else 0
        for name in self.ROLES:
            if name not in ps.__dict__:
                ps.register_auxiliary_attribute(name, "real")
        self.ps = ps
        self.n = ps.n
        dev = ps.mass.device
        self.scratch = torch.empty(2 * max(self.n, 1), dtype=torch.float64,
                                   device=dev)
        self.ctl = torch.zeros(6, dtype=torch.float64, device=dev)
        self.status = torch.zeros(1, dtype=torch.int32, device=dev)
        self.arrays = _native.ptr_array([ps.__dict__[r].data_ptr()
                                         for r in self.ROLES])
        self.fields = _native.ptr_array([ps.__dict__[f].data_ptr()
                                         for f in self.FIELDS])
        self.outs = _native.ptr_array([ps.__dict__[f].data_ptr()
                                       for f in self.OUT])
        self.capacity = 0

    def set_control(self, t_end):
        self.ctl.copy_(torch.tensor([float(type(self.ps).t_curr),
                                     float(t_end), float(self.integ.eta),
                                     0.0, 0.0, 0.0], dtype=torch.float64))

    def probe_capacity(self):
        """Near pairs of a pair drift of the current state
        (tupan_sakura_ex, two-pass form, both flags; the far/near split does
        not depend on dt; the outputs are scratch)."""
        import ctypes
        from ..lib.hip_backend import current_stream_handle
        most = 0
        for flag in (-1, 1):
            npairs = ctypes.c_int(-1)
            rc = self.lib.tupan_sakura_ex(
                self.n, self.fields, self.n, self.fields,
                abs(self.integ.eta) / 2, flag, self.outs, 2, None, None, None,
                ctypes.byref(npairs), self.status.data_ptr(),
                current_stream_handle())
            self.native.check(rc, "tupan_sakura_ex")
            most = max(most, npairs.value)
        self.status.zero_()
        self.capacity = max(4 * most, 4096)

    def launch(self):
        from ..lib.hip_backend import current_stream_handle
        rc = self.lib.tupan_sakura_step(
            self.adaptive, self.mode, self.n, self.arrays,
            self.scratch.data_ptr(), self.ctl.data_ptr(),
            float(self.integ.eta), self.capacity, None,
            self.status.data_ptr(), current_stream_handle())
        self.native.check(rc, "tupan_sakura_step")

    def read(self):
        """(device time, whether a near-pair list overflowed); raises on a
        failed step size or a Kepler failure."""
        c = self.ctl.cpu().tolist()
        st = int(self.status.item())
        if int(c[4]):
            raise self.native.TupanHipError(
                "tupan_sakura_step: %s" % self.STATUS.get(int(c[4]), c[4]))
        if st & 1:
            raise self.native.TupanHipError(
                "tupan_sakura_step: a pair exhausted the Kepler solver's "
                "bounded step subdivision")
        return c[0], bool(st & 2)

    def snapshot(self):
        snap = {k: self.ps.__dict__[k].clone() for k in self.STATE}
        snap["_ctl"] = self.ctl.clone()
        return snap

    def restore(self, snap):
        for k, v in snap.items():
            (self.ctl if k == "_ctl" else self.ps.__dict__[k]).copy_(v)

    def run(self, fn):
        """fn() (one launch or one replay) until no near-pair list overflows,
        doubling the capacity after each overflow; returns the new time and
        whether the capacity changed."""
        grown = False
        while True:
            snap = self.snapshot()
            self.status.zero_()
            fn()
            t, overflow = self.read()
            if not overflow:
                type(self.ps).t_curr = t
                return t, grown
            self.restore(snap)
            self.capacity *= 2
            grown = True
            fn = self.launch

    def capture(self, nsteps):
        graph = torch.cuda.CUDAGraph()
        side = torch.cuda.Stream()
        side.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(side):
            with torch.cuda.graph(graph, stream=side):
                for _ in range(nsteps):
                    self.launch()
        torch.cuda.current_stream().wait_stream(side)
        return graph


class Sakura(Base):
    PROVIDED_METHODS = ["sakura", "asakura"]

    def __init__(self, eta, time, ps, method, device_steps=False, **kwargs):
        super(Sakura, self).__init__(eta, time, ps, **kwargs)
        self.method = method
        self.device_steps = device_steps
        self._dev = None

    def _device(self):
        if self._dev is None or self._dev.ps is not self.ps:
            self._dev = DeviceSakura(self)
        return self._dev

    def evolve_step(self, t_end):
        """integrator/__init__.py:80-99; with device_steps one device step
        (and one read of the new time)."""
        if not self.device_steps:
            return super(Sakura, self).evolve_step(t_end)
        self.no_dumpper("device_steps")
        if not self.is_initialized:
            self.initialize(t_end)
        dev = self._device()
        dev.set_control(t_end)
        if dev.capacity == 0:
            dev.probe_capacity()
        dev.run(dev.launch)
        if self.reporter:
            self.reporter.diagnostic_report(self.ps)

    def evolve(self, t_end, steps_per_graph=16):
        """Advance to t_end by device steps, replaying a HIP graph of
        steps_per_graph steps (steps past t_end inside the last replay change
        nothing); a replay whose near-pair list outgrew its capacity is
        undone and run again with twice the capacity, and the graph is
        captured again.  Returns the number of replays."""
        if not self.device_steps:
            raise RuntimeError("evolve needs device_steps=True")
        self.no_dumpper("device_steps")
        if steps_per_graph < 1:
            raise ValueError("steps_per_graph must be at least 1")
        if not self.is_initialized:
            self.initialize(t_end)
        dev = self._device()
        dev.set_control(t_end)
        if dev.capacity == 0:
            dev.probe_capacity()
        t, _ = dev.run(dev.launch)        # eager first step: sizes workspaces
        replays = 0
        graph = None
        while (t < t_end) if self.eta > 0 else (t > t_end):
            if graph is None:
                graph = dev.capture(steps_per_graph)
            t_new, grown = dev.run(graph.replay)
            replays += 1
            if grown:
                # the overflowing replay was undone and one step run eagerly
                # with the larger list; the graph still holds the old one
                graph = None
            if t_new == t:
                raise dev.native.TupanHipError(
                    "device Sakura steps made no progress at t = %r" % t)
            t = t_new
        if self.reporter:
            self.reporter.diagnostic_report(self.ps)
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
        self.record_worldline(ps, tau)
        return ps
