"""Hermite predictor-corrector integrators of orders 2, 4, 6, 8 in P(EC)^n
form (tupan/integrator/hermite.py:20-410).

The predictor and corrector polynomials, their operation order, and the
shared adaptive step ('a' methods: Tstep then the power-of-two block step of
Base.get_min_block_tstep) are the reference's.  Force evaluations are the
extension singletons (Acc for order 2, AccJerk for 4, AccJerk + SnapCrackle
for 6 and 8); everything else is elementwise on the device arrays.

``Hermite(..., device_steps=True)`` hands each step to the device
(``DeviceSteps``, libtupan_hip: tupan_hermite_step): the step size, the
predictor, the corrector and the bookkeeping run as kernels that read the
step size from device memory, so ``evolve(t_end, steps_per_graph=k)`` can
record k steps once as a HIP graph and replay it, with one host
synchronisation per replay.
"""
import torch

from . import Base


class H2(object):
    @staticmethod
    def epredict(ps, tau):
        ps0 = ps.copy()
        ps0.set_acc(ps0)
        ps1 = ps
        ps1.rx += (ps0.ax * tau / 2 + ps0.vx) * tau
        ps1.ry += (ps0.ay * tau / 2 + ps0.vy) * tau
        ps1.rz += (ps0.az * tau / 2 + ps0.vz) * tau
        ps1.vx += ps0.ax * tau
        ps1.vy += ps0.ay * tau
        ps1.vz += ps0.az * tau
        return ps1, ps0

    @staticmethod
    def ecorrect(ps1, ps0, tau):
        ps1.set_acc(ps1)
        ps1.vx[...] = ((ps0.ax + ps1.ax) * tau / 2 + ps0.vx)
        ps1.vy[...] = ((ps0.ay + ps1.ay) * tau / 2 + ps0.vy)
        ps1.vz[...] = ((ps0.az + ps1.az) * tau / 2 + ps0.vz)
        ps1.rx[...] = ((ps0.vx + ps1.vx) * tau / 2 + ps0.rx)
        ps1.ry[...] = ((ps0.vy + ps1.vy) * tau / 2 + ps0.ry)
        ps1.rz[...] = ((ps0.vz + ps1.vz) * tau / 2 + ps0.rz)
        return ps1

    @classmethod
    def epec(cls, n, ps, tau):
        (ps1, ps0) = cls.epredict(ps, tau)
        for _ in range(n):
            ps1 = cls.ecorrect(ps1, ps0, tau)
        return ps1


class H4(H2):
    @staticmethod
    def epredict(ps, tau):
        ps0 = ps.copy()
        ps0.set_acc_jerk(ps0)
        ps1 = ps
        ps1.rx += ((ps0.jx * tau / 3 + ps0.ax) * tau / 2 + ps0.vx) * tau
        ps1.ry += ((ps0.jy * tau / 3 + ps0.ay) * tau / 2 + ps0.vy) * tau
        ps1.rz += ((ps0.jz * tau / 3 + ps0.az) * tau / 2 + ps0.vz) * tau
        ps1.vx += (ps0.jx * tau / 2 + ps0.ax) * tau
        ps1.vy += (ps0.jy * tau / 2 + ps0.ay) * tau
        ps1.vz += (ps0.jz * tau / 2 + ps0.az) * tau
        return ps1, ps0

    @staticmethod
    def ecorrect(ps1, ps0, tau):
        ps1.set_acc_jerk(ps1)
        ps1.vx[...] = (((ps0.jx - ps1.jx) * tau / 6
                        + (ps0.ax + ps1.ax)) * tau / 2 + ps0.vx)
        ps1.vy[...] = (((ps0.jy - ps1.jy) * tau / 6
                        + (ps0.ay + ps1.ay)) * tau / 2 + ps0.vy)
        ps1.vz[...] = (((ps0.jz - ps1.jz) * tau / 6
                        + (ps0.az + ps1.az)) * tau / 2 + ps0.vz)
        ps1.rx[...] = (((ps0.ax - ps1.ax) * tau / 6
                        + (ps0.vx + ps1.vx)) * tau / 2 + ps0.rx)
        ps1.ry[...] = (((ps0.ay - ps1.ay) * tau / 6
                        + (ps0.vy + ps1.vy)) * tau / 2 + ps0.ry)
        ps1.rz[...] = (((ps0.az - ps1.az) * tau / 6
                        + (ps0.vz + ps1.vz)) * tau / 2 + ps0.rz)
        return ps1


class H6(H4):
    @staticmethod
    def epredict(ps, tau):
        ps0 = ps.copy()
        ps0.set_acc_jerk(ps0)
        ps0.set_snap_crackle(ps0)
        ps1 = ps
        ps1.rx += (((ps0.sx * tau / 4 + ps0.jx) * tau / 3
                    + ps0.ax) * tau / 2 + ps0.vx) * tau
        ps1.ry += (((ps0.sy * tau / 4 + ps0.jy) * tau / 3
                    + ps0.ay) * tau / 2 + ps0.vy) * tau
        ps1.rz += (((ps0.sz * tau / 4 + ps0.jz) * tau / 3
                    + ps0.az) * tau / 2 + ps0.vz) * tau
        ps1.vx += ((ps0.sx * tau / 3 + ps0.jx) * tau / 2 + ps0.ax) * tau
        ps1.vy += ((ps0.sy * tau / 3 + ps0.jy) * tau / 2 + ps0.ay) * tau
        ps1.vz += ((ps0.sz * tau / 3 + ps0.jz) * tau / 2 + ps0.az) * tau
        return ps1, ps0

    @staticmethod
    def ecorrect(ps1, ps0, tau):
        ps1.set_acc_jerk(ps1)
        ps1.set_snap_crackle(ps1)
        ps1.vx[...] = ((((ps0.sx + ps1.sx) * tau / 12
                         + (ps0.jx - ps1.jx)) * tau / 5
                        + (ps0.ax + ps1.ax)) * tau / 2 + ps0.vx)
        ps1.vy[...] = ((((ps0.sy + ps1.sy) * tau / 12
                         + (ps0.jy - ps1.jy)) * tau / 5
                        + (ps0.ay + ps1.ay)) * tau / 2 + ps0.vy)
        ps1.vz[...] = ((((ps0.sz + ps1.sz) * tau / 12
                         + (ps0.jz - ps1.jz)) * tau / 5
                        + (ps0.az + ps1.az)) * tau / 2 + ps0.vz)
        ps1.rx[...] = ((((ps0.jx + ps1.jx) * tau / 12
                         + (ps0.ax - ps1.ax)) * tau / 5
                        + (ps0.vx + ps1.vx)) * tau / 2 + ps0.rx)
        ps1.ry[...] = ((((ps0.jy + ps1.jy) * tau / 12
                         + (ps0.ay - ps1.ay)) * tau / 5
                        + (ps0.vy + ps1.vy)) * tau / 2 + ps0.ry)
        ps1.rz[...] = ((((ps0.jz + ps1.jz) * tau / 12
                         + (ps0.az - ps1.az)) * tau / 5
                        + (ps0.vz + ps1.vz)) * tau / 2 + ps0.rz)
        return ps1


class H8(H6):
    @staticmethod
    def epredict(ps, tau):
        ps0 = ps.copy()
        ps0.set_acc_jerk(ps0)
        ps0.set_snap_crackle(ps0)
        ps1 = ps
        ps1.rx += ((((ps0.cx * tau / 5 + ps0.sx) * tau / 4 + ps0.jx) * tau / 3
                    + ps0.ax) * tau / 2 + ps0.vx) * tau
        ps1.ry += ((((ps0.cy * tau / 5 + ps0.sy) * tau / 4 + ps0.jy) * tau / 3
                    + ps0.ay) * tau / 2 + ps0.vy) * tau
        ps1.rz += ((((ps0.cz * tau / 5 + ps0.sz) * tau / 4 + ps0.jz) * tau / 3
                    + ps0.az) * tau / 2 + ps0.vz) * tau
        ps1.vx += (((ps0.cx * tau / 4 + ps0.sx) * tau / 3 + ps0.jx) * tau / 2
                   + ps0.ax) * tau
        ps1.vy += (((ps0.cy * tau / 4 + ps0.sy) * tau / 3 + ps0.jy) * tau / 2
                   + ps0.ay) * tau
        ps1.vz += (((ps0.cz * tau / 4 + ps0.sz) * tau / 3 + ps0.jz) * tau / 2
                   + ps0.az) * tau
        return ps1, ps0

    @staticmethod
    def ecorrect(ps1, ps0, tau):
        ps1.set_acc_jerk(ps1)
        ps1.set_snap_crackle(ps1)
        ps1.vx[...] = (((((ps0.cx - ps1.cx) * tau / 20
                          + (ps0.sx + ps1.sx)) * tau / 3
                         + 3 * (ps0.jx - ps1.jx)) * tau / 14
                        + (ps0.ax + ps1.ax)) * tau / 2 + ps0.vx)
        ps1.vy[...] = (((((ps0.cy - ps1.cy) * tau / 20
                          + (ps0.sy + ps1.sy)) * tau / 3
                         + 3 * (ps0.jy - ps1.jy)) * tau / 14
                        + (ps0.ay + ps1.ay)) * tau / 2 + ps0.vy)
        ps1.vz[...] = (((((ps0.cz - ps1.cz) * tau / 20
                          + (ps0.sz + ps1.sz)) * tau / 3
                         + 3 * (ps0.jz - ps1.jz)) * tau / 14
                        + (ps0.az + ps1.az)) * tau / 2 + ps0.vz)
        ps1.rx[...] = (((((ps0.sx - ps1.sx) * tau / 20
                          + (ps0.jx + ps1.jx)) * tau / 3
                         + 3 * (ps0.ax - ps1.ax)) * tau / 14
                        + (ps0.vx + ps1.vx)) * tau / 2 + ps0.rx)
        ps1.ry[...] = (((((ps0.sy - ps1.sy) * tau / 20
                          + (ps0.jy + ps1.jy)) * tau / 3
                         + 3 * (ps0.ay - ps1.ay)) * tau / 14
                        + (ps0.vy + ps1.vy)) * tau / 2 + ps0.ry)
        ps1.rz[...] = (((((ps0.sz - ps1.sz) * tau / 20
                          + (ps0.jz + ps1.jz)) * tau / 3
                         + 3 * (ps0.az - ps1.az)) * tau / 14
                        + (ps0.vz + ps1.vz)) * tau / 2 + ps0.rz)
        return ps1


_ORDERS = {2: H2, 4: H4, 6: H6, 8: H8}


class DeviceSteps(object):
    """Hermite steps of one integrator run by tupan_hermite_step: the
    particle arrays stay in place, a control block (t_curr, t_end, eta, tau)
    and a 32 n scratch array live on the device."""

    ROLES = ("mass", "eps2", "rx", "ry", "rz", "vx", "vy", "vz",
             "ax", "ay", "az", "jx", "jy", "jz", "sx", "sy", "sz",
             "cx", "cy", "cz", "tstep", "tstepij", "time", "nstep")
    SNAP_CRACKLE = ("sx", "sy", "sz", "cx", "cy", "cz")

    def __init__(self, integ):
        from .. import _native
        from ..lib import extensions
        from ..lib.hip_backend import HIPKernel
        ps = integ.ps
        k = getattr(extensions.acc_jerk, "kernel", None)
        if not isinstance(k, HIPKernel) or k.prec != "float64":
            raise RuntimeError("device Hermite steps need the HIP backend "
                               "in float64")
        if not ps.mass.is_cuda:
            raise RuntimeError("device Hermite steps need a device-resident "
                               "particle system")
        self.lib = _native.lib()
        self.native = _native
        self.mode = _native.MODES[k.mode]
        self.integ = integ
        self.order = integ.order
        self.adaptive = 1 if integ.method.startswith("a") else 0
        # s and c (snap, crackle) only for orders 6 and 8
        used = [r for r in self.ROLES
                if self.order >= 6 or r not in self.SNAP_CRACKLE]
        for name in used:
            if name not in ps.__dict__:
                ps.register_auxiliary_attribute(name, "real")
        self.ps = ps
        self.n = ps.n
        self.scratch = torch.empty(32 * max(self.n, 1), dtype=torch.float64,
                                   device=ps.mass.device)
        self.ctl = torch.zeros(5, dtype=torch.float64, device=ps.mass.device)
        ptrs = [ps.__dict__[r].data_ptr() if r in ps.__dict__ else None
                for r in self.ROLES]
        self.arrays = _native.ptr_array(ptrs)

    STATUS = {1: "non-finite time, t_end or eta",
              2: "Tstep gave no positive finite time step",
              3: "no power-of-two block step aligned with the time"}

    def set_control(self, t_end):
        self.ctl.copy_(torch.tensor([float(type(self.ps).t_curr), float(t_end),
                                     float(self.integ.eta), 0.0, 0.0],
                                    dtype=torch.float64))

    def launch(self):
        from ..lib.hip_backend import current_stream_handle
        rc = self.lib.tupan_hermite_step(
            self.order, self.adaptive, self.mode, self.n, self.arrays,
            self.scratch.data_ptr(), self.ctl.data_ptr(),
            float(self.integ.eta), current_stream_handle())
        self.native.check(rc, "tupan_hermite_step")

    def read_time(self):
        """The device time (one synchronisation); raises if a step failed."""
        c = self.ctl.cpu().tolist()
        status = int(c[4])
        if status:
            raise self.native.TupanHipError(
                "tupan_hermite_step: %s" % self.STATUS.get(status, status))
        type(self.ps).t_curr = c[0]
        return c[0]

    def capture(self, nsteps):
        """HIP graph of nsteps device steps (the arrays, scratch and control
        block are fixed in it)."""
        graph = torch.cuda.CUDAGraph()
        side = torch.cuda.Stream()
        side.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(side):
            with torch.cuda.graph(graph, stream=side):
                for _ in range(nsteps):
                    self.launch()
        torch.cuda.current_stream().wait_stream(side)
        return graph


class Hermite(Base):
    PROVIDED_METHODS = ["hermite2", "ahermite2",
                        "hermite4", "ahermite4",
                        "hermite6", "ahermite6",
                        "hermite8", "ahermite8"]

    def __init__(self, eta, time, ps, method, device_steps=False, **kwargs):
        super(Hermite, self).__init__(eta, time, ps, **kwargs)
        self.method = method
        self.order = int(method[-1])
        self.scheme = _ORDERS[self.order]
        self.device_steps = device_steps
        self._dev = None

    def _device(self):
        if self._dev is None or self._dev.ps is not self.ps:
            self._dev = DeviceSteps(self)
        return self._dev

    def evolve_step(self, t_end):
        """integrator/__init__.py:80-99; with device_steps one device step
        (and one read of the new time)."""
        if not self.device_steps:
            return super(Hermite, self).evolve_step(t_end)
        self.no_dumpper("device_steps")
        if not self.is_initialized:
            self.initialize(t_end)
        dev = self._device()
        dev.set_control(t_end)
        dev.launch()
        dev.read_time()
        if self.reporter:
            self.reporter.diagnostic_report(self.ps)

    def evolve(self, t_end, steps_per_graph=32):
        """Advance to t_end by device steps, replaying a HIP graph of
        steps_per_graph steps; steps past t_end inside the last replay change
        nothing.  Returns the number of replays."""
        if not self.device_steps:
            raise RuntimeError("evolve needs device_steps=True")
        self.no_dumpper("device_steps")
        if steps_per_graph < 1:
            raise ValueError("steps_per_graph must be at least 1")
        if not self.is_initialized:
            self.initialize(t_end)
        dev = self._device()
        dev.set_control(t_end)
        dev.launch()                      # eager first step: sizes workspaces
        replays = 0
        graph = None
        t = dev.read_time()
        while (t < t_end) if self.eta > 0 else (t > t_end):
            if graph is None:
                graph = dev.capture(steps_per_graph)
            graph.replay()
            replays += 1
            t_new = dev.read_time()
            if t_new == t:
                raise self.native.TupanHipError(
                    "device Hermite steps made no progress at t = %r" % t)
            t = t_new
        if self.reporter:
            self.reporter.diagnostic_report(self.ps)
        return replays

    def get_hermite_tstep(self, ps, eta, tau):
        ps.set_tstep(ps, eta)
        return self.get_min_block_tstep(ps, tau)

    def epec(self, n, ps, tau):
        return self.scheme.epec(n, ps, tau)

    def do_step(self, ps, tau):
        """hermite.py:390-410."""
        if "ahermite" in self.method:
            tau = self.get_hermite_tstep(ps, self.eta, tau)
        ps = self.epec(2, ps, tau)
        type(ps).t_curr += tau
        ps.tstep[...] = tau
        ps.time += tau
        ps.nstep += 1
        self.record_worldline(ps, tau)
        return ps
