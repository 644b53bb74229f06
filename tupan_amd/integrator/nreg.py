"""NREG: time-transformed leapfrog with the potential and kinetic energies as
step functions (tupan/integrator/nreg.py:22-175).

A position stage advances every particle with its velocity relative to the
centre of mass over dt = h / W (W = -U at start, then K - E0) through the
nreg_Xkernel sums (sum_j m_j r_ij(t + dt) / M); a velocity stage kicks over
dt = h / U with the accelerations at the drifted separations through the
nreg_Vkernel sums.  'anreg' is X(h/2W) V(h/U) X(h/2W) on half the step;
'nreg' composes two of them with the step-size function S updated in
between.  The physical time advances by the X stages' dt (t_curr).  The
system-wide sums U = sum_i u_i / 2 and K = sum_i k_i / (4 M) are exactly
rounded (math.fsum of the device values).
"""
import logging

from . import Base
from ..lib import extensions
from ..particles.system import fsum

logger = logging.getLogger(__name__)


def nreg_x(ps, dt):
    """nreg.py:22-43."""
    mtot = ps.total_mass
    extensions.nreg_x.calc(ps, ps, dt)
    ps.rx[...] = ps.mrx / mtot
    ps.ry[...] = ps.mry / mtot
    ps.rz[...] = ps.mrz / mtot
    type(ps).U = 0.5 * fsum(ps.u)
    type(ps).t_curr += dt
    return ps


def nreg_v(ps, dt):
    """nreg.py:46-75."""
    mtot = ps.total_mass
    extensions.nreg_v.calc(ps, ps, dt)
    ps.vx[...] = ps.mvx / mtot
    ps.vy[...] = ps.mvy / mtot
    ps.vz[...] = ps.mvz / mtot
    k = 0.25 * fsum(ps.mk) / mtot
    type(ps).W = k - ps.E0
    return ps


def anreg_step(ps, h):
    """nreg.py:78-87."""
    ps = nreg_x(ps, 0.5 * (h / ps.W))
    ps = nreg_v(ps, h / ps.U)
    ps = nreg_x(ps, 0.5 * (h / ps.W))
    return ps


def nreg_step(ps, h):
    """nreg.py:90-99."""
    ps = anreg_step(ps, 0.5 * (h * ps.S))
    type(ps).S = 1 / (2 / ps.W - 1 / ps.S)
    ps = anreg_step(ps, 0.5 * (h * ps.S))
    return ps


class NREG(Base):
    """nreg.py:102-175."""
    PROVIDED_METHODS = ["nreg", "anreg"]

    def __init__(self, eta, time, ps, method, **kwargs):
        super(NREG, self).__init__(eta, time, ps, **kwargs)
        self.method = method

    def initialize(self, t_end):
        ps = self.ps
        pe = ps.potential_energy
        type(ps).E0 = ps.kinetic_energy + pe
        type(ps).W = -pe
        type(ps).U = -pe
        type(ps).S = -pe
        if self.reporter:
            self.reporter.diagnostic_report(ps)
        if self.dumpper:
            self.dumpper.dump_worldline(ps)
        self.is_initialized = True

    def do_step(self, ps, tau):
        t0 = ps.t_curr
        if "anreg" in self.method:
            ps = anreg_step(ps, tau / 2)
        else:
            ps = nreg_step(ps, tau)
        dt = ps.t_curr - t0
        ps.tstep[...] = dt
        ps.time += tau
        ps.nstep += 1
        self.record_worldline(ps, tau)
        return ps
