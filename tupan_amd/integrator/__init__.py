"""Integrators (tupan/integrator/__init__.py:13-157).

``Base`` holds the block time-step logic shared by every method; the
``Integrator`` front end dispatches on the method name.  The methods run on
device-resident particle systems: every force evaluation goes through the
extension singletons (tupan_amd.lib.extensions), and the O(N) predictor,
corrector, drift and kick updates are elementwise device operations.
"""
import logging
import math
import sys

import numpy as np
import torch


class Base(object):
    def __init__(self, eta, time, ps, **kwargs):
        self.eta = eta
        type(ps).t_curr = time
        self.ps = ps
        self.is_initialized = False
        # integrator/__init__.py:23-36: a PN order switches on the PN
        # corrections (used by the SIA methods) and sets the module clight
        pn_order = kwargs.pop("pn_order", 0)
        clight = kwargs.pop("clight", None)
        type(ps).include_pn_corrections = False
        if pn_order > 0:
            if clight is None:
                raise TypeError(
                    "'clight' is not defined. Please set the speed of light "
                    "argument 'clight' when using 'pn_order' > 0.")
            from ..lib import extensions
            extensions.clight.pn_order = pn_order
            extensions.clight.clight = clight
            type(ps).include_pn_corrections = True
        self.reporter = kwargs.pop("reporter", None)
        self.viewer = kwargs.pop("viewer", None)
        self.dumpper = kwargs.pop("dumpper", None)
        self.dump_freq = kwargs.pop("dump_freq", 1)
        self.gl_freq = kwargs.pop("gl_freq", 1)
        if kwargs:
            msg = "{0}.__init__ received unexpected keyword arguments: {1}."
            raise TypeError(msg.format(type(self).__name__,
                                       ", ".join(kwargs.keys())))

    def get_base_tstep(self, t_end):
        """integrator/__init__.py:48-57."""
        ps = self.ps
        t_curr = ps.t_curr
        dt = min(abs(t_end) - abs(t_curr), abs(self.eta))
        dt = max(dt, abs(t_end) * (2 * sys.float_info.epsilon))
        return math.copysign(dt, self.eta)

    def get_min_block_tstep(self, ps, tau):
        """Power-of-two block step aligned with the current time
        (integrator/__init__.py:59-78).  As in the reference, t_next is
        formed once and the step halved until it divides t_next."""
        min_ts = ps.min_tstep()
        power = int(np.log2(min_ts) - 1)
        min_bts = 2.0**power
        t_curr = ps.t_curr
        t_next = t_curr + min_bts
        while t_next % min_bts != 0:
            min_bts /= 2
        min_bts = math.copysign(min_bts, tau)
        if abs(min_bts) > abs(tau):
            min_bts = tau
        return min_bts

    def evolve_step(self, t_end):
        """integrator/__init__.py:80-99: with a dumpper, the worldline
        samples that this step collects are written after it."""
        if not self.is_initialized:
            self.initialize(t_end)
        ps = self.ps
        self.wl = type(ps)(device=ps.device) if self.dumpper else None
        tau = self.get_base_tstep(t_end)
        ps = self.do_step(ps, tau)
        if self.reporter:
            self.reporter.diagnostic_report(ps)
        if self.dumpper:
            self.dumpper.dump_worldline(self.wl)
        self.ps = ps

    def initialize(self, t_end):
        """integrator/__init__.py:142-144 with hermite.py:312-328: report
        and dump the initial state."""
        if self.reporter:
            self.reporter.diagnostic_report(self.ps)
        if self.dumpper:
            self.dumpper.dump_worldline(self.ps)
        self.is_initialized = True

    def record_worldline(self, ps, tau):
        """hermite.py:402-405 (sakura.py:151-154, nreg.py:167-170,
        sia.py:1114-1117): with a dumpper, the particles of ``ps`` whose
        time is a multiple of dump_freq * tau join this step's worldline.
        fmod is exact, so its zeros are those of the reference's numpy
        ``%``."""
        if not self.dumpper:
            return
        slc = torch.fmod(ps.time, self.dump_freq * tau) == 0
        if bool(slc.any()):
            self.wl.append(ps[slc])

    def no_dumpper(self, what):
        """Device-driven steps do not collect worldlines."""
        if self.dumpper:
            raise ValueError("%s does not collect worldlines; use the "
                             "host-driven steps with a dumpper" % what)

    def finalize(self, t_end):
        pass


from . import hermite  # noqa: E402
from . import nreg  # noqa: E402
from . import sakura  # noqa: E402
from . import sia  # noqa: E402


class Integrator(object):
    """Method dispatch (integrator/__init__.py:108-157)."""
    PROVIDED_METHODS = []
    PROVIDED_METHODS.extend(sia.SIA.PROVIDED_METHODS)
    PROVIDED_METHODS.extend(sakura.Sakura.PROVIDED_METHODS)
    PROVIDED_METHODS.extend(hermite.Hermite.PROVIDED_METHODS)
    PROVIDED_METHODS.extend(nreg.NREG.PROVIDED_METHODS)

    def __init__(self, eta, time, ps, **kwargs):
        logger = logging.getLogger(__name__)
        method = kwargs.pop("method", None)
        self.integrator = None
        if method in sia.SIA.PROVIDED_METHODS:
            self.integrator = sia.SIA(eta, time, ps, method, **kwargs)
        elif method in sakura.Sakura.PROVIDED_METHODS:
            self.integrator = sakura.Sakura(eta, time, ps, method, **kwargs)
        elif method in hermite.Hermite.PROVIDED_METHODS:
            self.integrator = hermite.Hermite(eta, time, ps, method, **kwargs)
        elif method in nreg.NREG.PROVIDED_METHODS:
            self.integrator = nreg.NREG(eta, time, ps, method, **kwargs)
        else:
            logger.critical("Unexpected integration method: '%s'. Provided "
                            "methods: %s", method, str(self.PROVIDED_METHODS))
            raise ValueError("unexpected integration method %r" % (method,))

    def initialize(self, t_end):
        self.integrator.initialize(t_end)

    def finalize(self, t_end):
        self.integrator.finalize(t_end)

    def evolve_step(self, t_end):
        self.integrator.evolve_step(t_end)

    @property
    def time(self):
        return self.integrator.ps.t_curr

    @property
    def particle_system(self):
        return self.integrator.ps
