"""Symplectic integrators with hierarchical slow/fast splitting
(tupan/integrator/sia.py:24-352, 965-1123).

A scheme is a palindromic composition of drifts D(d_k tau) and kicks
K(k_k tau): DKD reads (d0 k0 d1 k1 ...) forwards to the middle and back,
KDK the same with the roles of the two coefficient lists exchanged
(sia.py:300-352 and the SIA22..SIA69 classes).  The coefficient tables are
the published ones the reference uses: Yoshida, Phys. Lett. A 150 (1990)
(SIA21, SIA43, SIA67); Omelyan, Mryglod & Folk, Comput. Phys. Comm. 151
(2003) (SIA22, SIA44, SIA45); Blanes & Moan, J. Comp. Appl. Math. 142 (2002)
(SIA46); Kahan & Li, Math. Comput. 66 (1997) (SIA69).

Methods 'siaNNx.dkd' / 'siaNNx.kdk' with x = 's' (shared fixed step),
'a' (shared adaptive block step) or 'h' (hierarchical): each level computes
Tstep on its subset, splits it into slow (|tstep| > tau) and fast parts,
advances the slow part with the scheme, recurses on the fast part with the
scheme's drift substeps, and couples them by slow<->fast kicks
(sia.py:1089-1123).

Subsets.  The reference splits with copies (``ps[mask]``) and joins by
appending fast after slow, so after every level the particle order is
[slow | fast], recursively.  By default this driver produces that order in
place instead: ``partition`` stably permutes the level's range of the device
arrays to [slow | fast] and hands out slice views of the two parts, so every
kernel call sees the same particles in the same order as the reference's
copies, no subset is ever copied out, and the join is a no-op.
``SIA(..., inplace=False)`` runs the reference's copy/append form.

Post-Newtonian corrections.  With ``pn_order`` and ``clight`` given, kicks
add the PN pair accelerations through an auxiliary velocity w (kick_pn and
the PN branch of sf_kick, sia.py:104-159, 241-294) and keep the PN work,
impulse and torque in pn_ke, pn_mv, pn_am (drifts move pn_mr), as the
reference does.  As there, a subset of exactly two particles raises
NotImplementedError (FewBody's Kepler solver has no PN terms), so with PN
the 's' and 'a' variants, which never split off pairs, are the usable ones.

Small subsets.  With the HIP backend in FP64 or split-FP32 mode and no PN
corrections, the 'h' variants hand every subset of at most SUBTREE_MAX particles to the device
(libtupan_hip: tupan_sia_subtree), which runs that subset's whole recursion
in one launch with the same operations in the same order, so the deep
levels -- most of the visits, each a few tiny kernels -- cost one launch per
subset instead of host work per visit.  ``SIA(..., device_subtree=False)``
keeps every level on the host.
"""
import ctypes
import logging

import torch

from . import Base
from .fewbody import FewBody

SUBTREE_MAX = 64
SUBTREE_ROLES = ("mass", "eps2", "rx", "ry", "rz", "vx", "vy", "vz", "time",
                 "nstep", "tstep", "tstepij", "ax", "ay", "az")

logger = logging.getLogger(__name__)

COEFS = {
    "sia21": ([1.0],
              [0.5]),
    "sia22": ([0.5],
              [0.1931833275037836, 0.6136333449924328]),
    "sia43": ([1.3512071919596575, -1.7024143839193150],
              [0.6756035959798288, -0.17560359597982877]),
    "sia44": ([0.7123418310626056, -0.21234183106260562],
              [0.1786178958448091, -0.06626458266981843, 0.7752933736500186]),
    "sia45": ([-0.0844296195070715, 0.354900057157426, 0.459059124699291],
              [0.2750081212332419, -0.1347950099106792, 0.35978688867743724]),
    "sia46": ([0.209515106613362, -0.143851773179818, 0.434336666566456],
              [0.0792036964311957, 0.353172906049774, -0.0420650803577195,
               0.21937695575349958]),
    "sia67": ([0.7845136104775573, 0.23557321335935813, -1.177679984178871,
               1.3151863206839112],
              [0.39225680523877865, 0.5100434119184577, -0.47105338540975644,
               0.06875316825252015]),
    "sia69": ([0.39103020330868477, 0.334037289611136, -0.7062272811875614,
               0.08187754964805945, 0.7985644772393624],
              [0.19551510165434238, 0.3625337464599104, -0.1860949957882127,
               -0.31217486576975095, 0.44022101344371095]),
}


def palindrome(first, second):
    """Alternate first[0], second[0], first[1], ... then mirror; returns
    [(which, coef)] with which 0 for `first`, 1 for `second`."""
    seq = []
    i = 0
    while i < len(first) or i < len(second):
        if i < len(first):
            seq.append((0, first[i]))
        if i < len(second):
            seq.append((1, second[i]))
        i += 1
    return seq + seq[-2::-1]


# -- operators (sia.py:24-212) ------------------------------------------------

def split(ps, condition):
    if ps.n <= 2:
        return ps, ps[torch.zeros(ps.n, dtype=torch.bool, device=ps.device)]
    slow = ps[condition]
    fast = ps[~condition]
    if slow.n + fast.n != ps.n:
        logger.error("slow.n + fast.n != ps.n: %d, %d, %d.",
                     slow.n, fast.n, ps.n)
    return slow, fast


def partition(ps, condition):
    """In-place form of ``split``: stably permute ps's range of every array
    to [slow | fast] and return slice views of the two parts.  ``ps`` must
    be a system or a slice view of one whose arrays are all registered."""
    n = ps.n
    if n <= 2:
        return ps, ps[n:n]
    slow_idx = torch.nonzero(condition).squeeze(1)
    ns = int(slow_idx.shape[0])
    if 0 < ns < n:
        perm = torch.cat([slow_idx, torch.nonzero(~condition).squeeze(1)])
        for t in ps.arrays().values():
            t.copy_(t.index_select(0, perm))
    return ps[0:ns], ps[ns:n]


def join(slow, fast):
    if not fast.n:
        return slow
    if not slow.n:
        return fast
    ps = slow
    ps.append(fast)
    return ps


def drift(ips, tau):
    """sia.py:162-171 with drift_n and drift_pn (sia.py:64-101)."""
    ips.rx += ips.vx * tau
    ips.ry += ips.vy * tau
    ips.rz += ips.vz * tau
    if ips.include_pn_corrections:
        ips.pn_drift_com_r(tau)
    return ips


def kick_n(ips, tau):
    ips.vx += ips.ax * tau
    ips.vy += ips.ay * tau
    ips.vz += ips.az * tau
    return ips


def _pn_half(ps, tau):
    """v += (a tau + w) / 2."""
    ps.vx += (ps.ax * tau + ps.wx) / 2
    ps.vy += (ps.ay * tau + ps.wy) / 2
    ps.vz += (ps.az * tau + ps.wz) / 2


def _pn_book(ps, tau):
    """PN bookkeeping of one kick, then w <- 2 a_PN tau - w."""
    ps.pn_kick_ke(tau)
    ps.pn_kick_lmom(tau)
    ps.pn_kick_amom(tau)
    ps.wx[...] = 2 * ps.pnax * tau - ps.wx
    ps.wy[...] = 2 * ps.pnay * tau - ps.wy
    ps.wz[...] = 2 * ps.pnaz * tau - ps.wz


def kick_pn(ips, tau):
    """sia.py:104-159: the velocity-dependent PN acceleration enters through
    the auxiliary velocity w, evaluated once between two half kicks."""
    _pn_half(ips, tau)
    ips.set_pnacc(ips)
    _pn_book(ips, tau)
    _pn_half(ips, tau)
    return ips


def kick(ips, tau):
    """sia.py:177-186."""
    ips.set_acc(ips)
    if ips.include_pn_corrections:
        return kick_pn(ips, tau)
    return kick_n(ips, tau)


def sf_drift(slow, fast, tau, evolve, recurse, bridge):
    slow = evolve(slow, tau)
    fast = recurse(fast, tau, evolve, bridge)
    return slow, fast


def sf_kick(slow, fast, tau):
    """sia.py:201-294."""
    if slow.n and fast.n:
        slow.set_acc(fast)
        fast.set_acc(slow)
        if slow.include_pn_corrections:
            _pn_half(slow, tau)
            _pn_half(fast, tau)
            slow.set_pnacc(fast)
            fast.set_pnacc(slow)
            _pn_book(slow, tau)
            _pn_book(fast, tau)
            _pn_half(slow, tau)
            _pn_half(fast, tau)
        else:
            slow = kick_n(slow, tau)
            fast = kick_n(fast, tau)
    return slow, fast


class Scheme(object):
    """One coefficient set with its DKD, KDK and slow/fast bridge."""

    def __init__(self, name):
        self.name = name
        self.k, self.d = COEFS[name]

    def dkd(self, ips, tau):
        if ips.n <= 2:
            return FewBody.evolve(ips, tau)
        for which, c in palindrome(self.d, self.k):
            ips = drift(ips, c * tau) if which == 0 else kick(ips, c * tau)
        return ips

    def kdk(self, ips, tau):
        if ips.n <= 2:
            return FewBody.evolve(ips, tau)
        # KDK exchanges the roles of the two lists (sia.py:325)
        d, k = self.k, self.d
        for which, c in palindrome(k, d):
            ips = kick(ips, c * tau) if which == 0 else drift(ips, c * tau)
        return ips

    def bridge_sf(self, slow, fast, tau, evolve, recurse):
        for which, c in palindrome(self.d, self.k):
            if which == 0:
                slow, fast = sf_drift(slow, fast, c * tau, evolve, recurse,
                                      self.bridge_sf)
            else:
                slow, fast = sf_kick(slow, fast, c * tau)
        return slow, fast


class SIA(Base):
    PROVIDED_METHODS = ["%s%s.%s" % (s, v, t)
                        for s in sorted(COEFS) for v in "sah"
                        for t in ("dkd", "kdk")]

    # outputs of the kernels the method calls, registered on the whole
    # system up front so that slice views share them; with PN corrections
    # also the auxiliary velocity w (sia.py:1002-1005) and the PN
    # bookkeeping arrays
    AUX = ("tstepij", "ax", "ay", "az")
    PN_AUX = ("wx", "wy", "wz", "pnax", "pnay", "pnaz", "pn_ke", "pn_mrx",
              "pn_mry", "pn_mrz", "pn_mvx", "pn_mvy", "pn_mvz", "pn_amx",
              "pn_amy", "pn_amz")

    def __init__(self, eta, time, ps, method, inplace=True,
                 device_subtree=True, device_steps=False, device_levels=False,
                 **kwargs):
        super(SIA, self).__init__(eta, time, ps, **kwargs)
        self.inplace = inplace
        self.device_subtree = device_subtree
        self.device_steps = device_steps
        self.device_levels = device_levels
        self._dev = None
        self._levels = None
        self.method = method
        self.scheme = Scheme(method[:5])
        variant = method[5]
        self.update_tstep = variant in "ah"
        self.shared_tstep = variant in "sa"
        self.evolve_set = self.scheme.dkd if method.endswith("dkd") \
            else self.scheme.kdk

    def _device(self):
        if self._dev is None or self._dev.ps is not self.ps:
            self._dev = DeviceSteps(self)
        return self._dev

    def _device_levels(self):
        if self._levels is None or not self._levels.valid_for(self.ps):
            if self._levels is not None:
                self._levels.close()
            self._levels = DeviceLevels(self)
        return self._levels

    def evolve_step(self, t_end):
        """integrator/__init__.py:80-99; with device_steps (the 's' and 'a'
        variants) one device step and one read of the new time; with
        device_levels (the 'h' variants) one device-scheduled step; a set of
        at most 64 particles, which the host driver advances in one subtree
        launch, takes the host driver's step."""
        if self.device_levels and not self.device_steps and \
                self.ps.n > SUBTREE_MAX:
            self.no_dumpper("device_levels")
            if not self.is_initialized:
                self.initialize(t_end)
            self._device_levels().step(t_end,
                                       graphs=self.device_levels != "eager")
            if self.reporter:
                self.reporter.diagnostic_report(self.ps)
            return
        if not self.device_steps:
            return super(SIA, self).evolve_step(t_end)
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
        """Advance to t_end by device steps ('s' and 'a' variants),
        replaying a HIP graph of steps_per_graph steps (steps past t_end
        inside the last replay change nothing), or by device-scheduled
        level steps ('h' variants with device_levels), each replaying the
        level graphs until the device reports the step done.  Returns the
        number of replays."""
        if self.device_levels and not self.device_steps:
            self.no_dumpper("device_levels")
            if steps_per_graph < 1:
                raise ValueError("steps_per_graph must be at least 1")
            if self.ps.n <= SUBTREE_MAX:
                # one subtree launch per step already
                while (self.ps.t_curr < t_end) if self.eta > 0 else \
                        (self.ps.t_curr > t_end):
                    self.evolve_step(t_end)
                return 0
            if not self.is_initialized:
                self.initialize(t_end)
            replays = self._device_levels().evolve(t_end)
            if self.reporter:
                self.reporter.diagnostic_report(self.ps)
            return replays
        if not self.device_steps:
            raise RuntimeError("evolve needs device_steps=True or "
                               "device_levels=True")
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
                raise dev.native.TupanHipError(
                    "device SIA steps made no progress at t = %r" % t)
            t = t_new
        if self.reporter:
            self.reporter.diagnostic_report(self.ps)
        return replays

    def do_step(self, ps, tau):
        names = self.AUX if self.inplace else ()
        if ps.include_pn_corrections:
            names = names + self.PN_AUX
        for name in names:
            if name not in ps.__dict__:
                ps.register_auxiliary_attribute(name, "real")
        # the one-launch subtree runs its levels on the device, where no
        # worldline samples are taken
        self._subtree = (None if self.dumpper
                         else _DeviceSubtree.make(self, ps))
        return self.recurse(ps, tau, self.evolve_set, self.scheme.bridge_sf)

    def recurse(self, ps, tau, evolve, bridge_sf):
        """sia.py:1089-1123."""
        if not ps.n:
            return ps
        sub = getattr(self, "_subtree", None)
        if (sub is not None and ps.n <= SUBTREE_MAX
                and evolve == self.evolve_set):
            return sub.run(ps, tau)
        flag = -1
        if self.update_tstep:
            flag = 1
            ps.set_tstep(ps, self.eta)
            if self.shared_tstep:
                tau = self.get_min_block_tstep(ps, tau)
        cond = abs(ps.tstep) > flag * abs(tau)
        slow, fast = (partition if self.inplace else split)(ps, cond)
        slow, fast = bridge_sf(slow, fast, tau, evolve, self.recurse)
        if fast.n == 0:
            type(ps).t_curr += tau
        if slow.n:
            slow.tstep[...] = tau
            slow.time += tau
            slow.nstep += 1
            self.record_worldline(slow, tau)
        return ps if self.inplace else join(slow, fast)


class DeviceSteps(object):
    """Shared-step SIA steps ('s' and 'a' variants) run by tupan_sia_step:
    every particle is slow at every step, so a step is the bridge's
    evolve(all, c_b tau) calls flattened into one drift/kick sequence with
    h = c_e * (c_b * tau); the step size lives in a device control block
    (t_curr, t_end, eta, tau, status) and a 5 n scratch array holds Acc and
    Tstep outputs until the step commits them."""

    ROLES = ("mass", "eps2", "rx", "ry", "rz", "vx", "vy", "vz", "ax", "ay",
             "az", "tstep", "tstepij", "time", "nstep")
    STATUS = {1: "non-finite time, t_end or eta",
              2: "Tstep gave no positive finite time step",
              3: "no power-of-two block step aligned with the time"}

    def __init__(self, sia):
        from .. import _native
        from ..lib import extensions
        from ..lib.hip_backend import HIPKernel
        ps = sia.ps
        if not sia.shared_tstep:
            raise RuntimeError("device SIA steps run the shared-step 's' and "
                               "'a' variants; %s is hierarchical" % sia.method)
        if ps.include_pn_corrections:
            raise RuntimeError("device SIA steps have no PN corrections")
        modes = []
        for ext in (extensions.acc, extensions.tstep):
            k = getattr(ext, "kernel", None)
            if not isinstance(k, HIPKernel) or k.prec != "float64":
                raise RuntimeError("device SIA steps need the HIP backend in "
                                   "float64")
            modes.append(_native.MODES[k.mode])
        if not ps.mass.is_cuda:
            raise RuntimeError("device SIA steps need a device-resident "
                               "particle system")
        if ps.n < 3:
            raise RuntimeError("device SIA steps need at least 3 particles "
                               "(1 or 2 evolve by FewBody)")
        for name in self.ROLES:
            if name not in ps.__dict__:
                ps.register_auxiliary_attribute(name, "real")
        self.lib = _native.lib()
        self.native = _native
        self.sia = sia
        self.ps = ps
        self.n = ps.n
        self.acc_mode, self.tstep_mode = modes
        self.adaptive = 0 if sia.method[5] == "s" else 1
        seq = palindrome(sia.scheme.d, sia.scheme.k)
        if sia.method.endswith("dkd"):
            ev = [(0 if w == 0 else 1, c) for w, c in seq]
        else:
            ev = [(1 if w == 0 else 0, c) for w, c in seq]
        ops, cb, ce = [], [], []
        for wb, c_b in seq:
            if wb == 0:          # sf_drift: evolve(slow, c_b tau); no fast set
                for we, c_e in ev:
                    ops.append(we)
                    cb.append(c_b)
                    ce.append(c_e)
            # wb == 1: sf_kick with an empty fast set does nothing
        self.nops = len(ops)
        self.ops = (ctypes.c_int * len(ops))(*ops)
        self.cb = (ctypes.c_double * len(cb))(*cb)
        self.ce = (ctypes.c_double * len(ce))(*ce)
        dev = ps.mass.device
        self.scratch = torch.empty(5 * self.n, dtype=torch.float64, device=dev)
        self.ctl = torch.zeros(5, dtype=torch.float64, device=dev)
        self.arrays = _native.ptr_array([ps.__dict__[r].data_ptr()
                                         for r in self.ROLES])

    def set_control(self, t_end):
        self.ctl.copy_(torch.tensor([float(type(self.ps).t_curr),
                                     float(t_end), float(self.sia.eta), 0.0,
                                     0.0], dtype=torch.float64))

    def launch(self):
        from ..lib.hip_backend import current_stream_handle
        rc = self.lib.tupan_sia_step(
            self.adaptive, self.acc_mode, self.tstep_mode, self.n,
            self.arrays, self.scratch.data_ptr(), self.ctl.data_ptr(),
            float(self.sia.eta), self.nops, ctypes.addressof(self.ops),
            ctypes.addressof(self.cb), ctypes.addressof(self.ce),
            current_stream_handle())
        self.native.check(rc, "tupan_sia_step")

    def read_time(self):
        """The device time (one synchronisation); raises if a step failed."""
        c = self.ctl.cpu().tolist()
        status = int(c[4])
        if status:
            raise self.native.TupanHipError(
                "tupan_sia_step: %s" % self.STATUS.get(status, status))
        type(self.ps).t_curr = c[0]
        return c[0]

    def capture(self, nsteps):
        """HIP graph of nsteps device steps."""
        graph = torch.cuda.CUDAGraph()
        side = torch.cuda.Stream()
        side.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(side):
            with torch.cuda.graph(graph, stream=side):
                for _ in range(nsteps):
                    self.launch()
        torch.cuda.current_stream().wait_stream(side)
        return graph


class DeviceLevels(object):
    """Hierarchical ('h') SIA steps walked on the device
    (tupan_sia_levels_*): the level visits, partitions, drifts, kicks and
    pairwise calls of the host driver, scheduled from device memory one
    operation per slot, so a step needs no host read of level sizes.  A step
    runs as slots in batches of `slots`, replayed from two captured HIP
    graphs (the step's start plus a batch, and a batch) or, with
    SIA(..., device_levels="eager"), enqueued call by call, until the device
    reports the step done."""

    STATUS = {1: "Kepler solver failure", 2: "more than 64 levels",
              3: "more than 2^20 level visits",
              4: "operation queue overflow", 16: "scheduler bound",
              32: "no valid step size"}

    def __init__(self, sia, slots=256):
        from .. import _native
        from ..lib import extensions
        from ..lib.hip_backend import HIPKernel
        ps = sia.ps
        if not (sia.inplace and sia.update_tstep and not sia.shared_tstep):
            raise RuntimeError("device levels run the hierarchical 'h' "
                               "variants in place")
        if ps.include_pn_corrections:
            raise RuntimeError("device levels have no PN corrections")
        modes = []
        for ext in (extensions.acc, extensions.tstep):
            k = getattr(ext, "kernel", None)
            if not isinstance(k, HIPKernel) or k.prec != "float64" or \
                    k.mode not in ("fp64", "split32"):
                raise RuntimeError("device levels need the HIP backend in "
                                   "float64, fp64 or split32 mode")
            modes.append(_native.MODES[k.mode])
        if ps.n <= SUBTREE_MAX:
            raise RuntimeError("device levels need more than %d particles"
                               % SUBTREE_MAX)
        for name in sia.AUX:
            if name not in ps.__dict__:
                ps.register_auxiliary_attribute(name, "real")
        arrays = ps.arrays()
        if not all(t.is_cuda and t.element_size() == 8
                   for t in arrays.values()):
            raise RuntimeError("device levels need device arrays of 8-byte "
                               "elements")
        if any(r not in arrays for r in SUBTREE_ROLES) or len(arrays) > 32:
            raise RuntimeError("device levels need the SIA arrays, at most "
                               "32 of them")
        self.lib = _native.lib()
        self.native = _native
        self.sia = sia
        self.ps = ps
        self.names = list(arrays)
        self.ptrs = [t.data_ptr() for t in arrays.values()]
        self.slots = slots
        roles = (ctypes.c_int * len(SUBTREE_ROLES))(
            *[self.names.index(r) for r in SUBTREE_ROLES])
        seq = palindrome(sia.scheme.d, sia.scheme.k)
        bop = (ctypes.c_int * len(seq))(*[w for w, _ in seq])
        bc = (ctypes.c_double * len(seq))(*[c for _, c in seq])
        if sia.method.endswith("dkd"):
            ev = [(0 if w == 0 else 1, c) for w, c in seq]
        else:
            ev = [(1 if w == 0 else 0, c) for w, c in seq]
        eop = (ctypes.c_int * len(ev))(*[w for w, _ in ev])
        ec = (ctypes.c_double * len(ev))(*[c for _, c in ev])
        self.plan = ctypes.c_void_p()
        rc = self.lib.tupan_sia_levels_create(
            ps.n, _native.ptr_array(self.ptrs), len(self.names),
            ctypes.addressof(roles), modes[0], modes[1], float(sia.eta),
            len(seq), ctypes.addressof(bop), ctypes.addressof(bc), len(ev),
            ctypes.addressof(eop), ctypes.addressof(ec),
            ctypes.byref(self.plan))
        _native.check(rc, "tupan_sia_levels_create")
        self.graphs = None

    def valid_for(self, ps):
        arrays = ps.arrays()
        return (ps is self.ps and list(arrays) == self.names
                and [t.data_ptr() for t in arrays.values()] == self.ptrs)

    def close(self):
        if self.plan:
            self.lib.tupan_sia_levels_destroy(self.plan)
            self.plan = ctypes.c_void_p()
        self.graphs = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_control(self, t_end, clear_pairs=True):
        rc = self.lib.tupan_sia_levels_set(
            self.plan, float(type(self.ps).t_curr), float(t_end),
            float(self.sia.eta), 1 if clear_pairs else 0)
        self.native.check(rc, "tupan_sia_levels_set")

    def enqueue(self, begin):
        from ..lib.hip_backend import current_stream_handle
        rc = self.lib.tupan_sia_levels_enqueue(
            self.plan, 1 if begin else 0, self.slots, current_stream_handle())
        self.native.check(rc, "tupan_sia_levels_enqueue")

    def state(self):
        """(t_curr, done, visits, (Tstep, Acc) interactions); raises if the
        walk failed.  One synchronisation.  Also keeps the step's level
        visits above 64 particles and slots that did work in
        last_visits and last_ops."""
        step = (ctypes.c_double * 5)()
        st = (ctypes.c_int * 4)()
        pairs = (ctypes.c_ulonglong * 2)()
        rc = self.lib.tupan_sia_levels_state(self.plan, step, st, pairs)
        self.native.check(rc, "tupan_sia_levels_state")
        if st[1]:
            bits = [m for b, m in sorted(self.STATUS.items()) if st[1] & b]
            raise self.native.TupanHipError(
                "device SIA levels failed (status %d: %s)"
                % (st[1], ", ".join(bits) or "?"))
        self.last_visits, self.last_ops = st[2], st[3]
        return step[0], bool(st[0]), st[2], (pairs[0], pairs[1])

    def _run(self, first, more):
        """One step: `first` enqueues or replays the start and a batch,
        `more` a further batch, until the device reports the step done."""
        from ..lib import extensions
        first()
        t, done, _, pairs = self.state()
        batches = 1
        while not done:
            more()
            t, done, _, pairs = self.state()
            batches += 1
        type(self.ps).t_curr = t
        if extensions.pair_counts is not None:
            extensions.pair_counts["Tstep"] += pairs[0]
            extensions.pair_counts["Acc"] += pairs[1]
        return batches

    def step(self, t_end, graphs=True):
        """One step: slots replayed from the two captured graphs (captured
        on first use), or enqueued eagerly.  Returns the replays or
        batches."""
        self.set_control(t_end)
        if not graphs:
            return self._run(lambda: self.enqueue(True),
                             lambda: self.enqueue(False))
        if self.graphs is None:
            self.graphs = self.capture()
        return self._run(self.graphs[0].replay, self.graphs[1].replay)

    def capture(self):
        graphs = []
        for begin in (True, False):
            g = torch.cuda.CUDAGraph()
            side = torch.cuda.Stream()
            side.wait_stream(torch.cuda.current_stream())
            with torch.cuda.stream(side):
                with torch.cuda.graph(g, stream=side):
                    self.enqueue(begin)
            torch.cuda.current_stream().wait_stream(side)
            graphs.append(g)
        return graphs

    def evolve(self, t_end):
        """Steps to t_end from the captured graphs; returns the number of
        graph replays."""
        if self.graphs is None:
            self.graphs = self.capture()
        g_first, g_more = self.graphs
        replays = 0
        t = type(self.ps).t_curr
        while (t < t_end) if self.sia.eta > 0 else (t > t_end):
            self.set_control(t_end)
            replays += self._run(g_first.replay, g_more.replay)
            t_new = type(self.ps).t_curr
            if t_new == t:
                raise self.native.TupanHipError(
                    "device SIA levels made no progress at t = %r" % t)
            t = t_new
        return replays


class _DeviceSubtree(object):
    """The host side of tupan_sia_subtree for one SIA object: the scheme's
    operator sequences, the arithmetic modes of the installed Acc and Tstep
    kernels, and small device buffers for the time, counters and status."""

    @classmethod
    def make(cls, sia, ps):
        from ..lib import extensions
        from ..lib.hip_backend import HIPKernel
        if not (sia.device_subtree and sia.inplace and sia.update_tstep
                and not sia.shared_tstep and not ps.include_pn_corrections):
            return None
        kernels = []
        for ext in (extensions.acc, extensions.tstep):
            k = getattr(ext, "kernel", None)
            if not isinstance(k, HIPKernel) or k.prec != "float64" or \
                    k.mode not in ("fp64", "split32"):
                return None
            kernels.append(k)
        arrays = ps.arrays()
        if not all(t.is_cuda and t.element_size() == 8
                   for t in arrays.values()):
            return None
        if any(r not in arrays for r in SUBTREE_ROLES) or len(arrays) > 32:
            return None
        return cls(sia, kernels[0].mode, kernels[1].mode, ps.device)

    def __init__(self, sia, acc_mode, tstep_mode, device):
        from .. import _native
        self.lib = _native.lib()
        self.sia = sia
        self.modes = (_native.MODES[acc_mode], _native.MODES[tstep_mode])
        seq = palindrome(sia.scheme.d, sia.scheme.k)
        self.bridge_op = (ctypes.c_int * len(seq))(*[w for w, _ in seq])
        self.bridge_c = (ctypes.c_double * len(seq))(*[c for _, c in seq])
        if sia.method.endswith("dkd"):
            ev = [(0 if w == 0 else 1, c) for w, c in seq]
        else:
            ev = [(1 if w == 0 else 0, c) for w, c in seq]
        self.evolve_op = (ctypes.c_int * len(ev))(*[w for w, _ in ev])
        self.evolve_c = (ctypes.c_double * len(ev))(*[c for _, c in ev])
        self.t_curr = torch.zeros(1, dtype=torch.float64, device=device)
        self.pairs = torch.zeros(2, dtype=torch.int64, device=device)
        self.status = torch.zeros(1, dtype=torch.int32, device=device)

    def run(self, ps, tau):
        from .. import _native
        from ..lib import extensions
        from ..lib.hip_backend import current_stream_handle
        arrays = ps.arrays()
        names = list(arrays)
        ptrs = _native.ptr_array([t.data_ptr() for t in arrays.values()])
        roles = (ctypes.c_int * len(SUBTREE_ROLES))(
            *[names.index(r) for r in SUBTREE_ROLES])
        self.t_curr.fill_(type(ps).t_curr)
        self.pairs.zero_()
        self.status.zero_()
        rc = self.lib.tupan_sia_subtree(
            ps.n, ptrs, len(names), ctypes.addressof(roles), self.modes[0],
            self.modes[1], self.sia.eta, tau, len(self.bridge_c),
            ctypes.addressof(self.bridge_op), ctypes.addressof(self.bridge_c),
            len(self.evolve_c), ctypes.addressof(self.evolve_op),
            ctypes.addressof(self.evolve_c), self.t_curr.data_ptr(),
            self.pairs.data_ptr(), self.status.data_ptr(),
            current_stream_handle())
        _native.check(rc, "tupan_sia_subtree")
        status = int(self.status.item())
        if status:
            raise _native.TupanHipError(
                "SIA subtree failed (status %d: 1 Kepler solver, 2 more than "
                "64 levels, 3 more than 65536 level visits)" % status)
        type(ps).t_curr = float(self.t_curr.item())
        if extensions.pair_counts is not None:
            tp, ap = self.pairs.tolist()
            extensions.pair_counts["Tstep"] += tp
            extensions.pair_counts["Acc"] += ap
        return ps
