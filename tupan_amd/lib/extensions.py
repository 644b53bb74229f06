"""High-level interfaces to the gravity kernels (tupan/lib/extensions.py).

Same classes, constructor signature ``Cls(backend, prec)``, methods
``set_args / run / get_result / calc`` and module-level singletons as the
reference (extensions.py:84-666).  ``calc(ips, jps, *scalars)`` evaluates the
kernel for the i-particles ``ips`` due to the j-particles ``jps`` and returns
the output arrays, which are auxiliary attributes of ``ips`` registered on
first use (zero-filled) and overwritten by each call.

Backends: ``"HIP"`` (libtupan_hip.so, tupan_amd.lib.hip_backend) is the
product backend and the default of the singletons.  As in the reference
(extensions.py:63-76), ``get_kernel`` raises ``ValueError`` for an unknown
backend name; a backend may also be given as a class with the reference
kernel-object protocol, constructed as ``cls(prec, name)``.

PNAcc (the post-Newtonian kernel) is not provided by this build; its
singleton raises ``NotImplementedError`` when used.
"""
import collections
import logging
import sys

from .utils import ctype

__all__ = ["Phi", "phi",
           "Acc", "acc",
           "AccJerk", "acc_jerk",
           "SnapCrackle", "snap_crackle",
           "Tstep", "tstep",
           "Sakura", "sakura",
           "Kepler", "kepler",
           "NREG_X", "nreg_x",
           "NREG_V", "nreg_v",
           "PNAcc", "pnacc", "Clight", "clight",
           "get_kernel", "install"]

logger = logging.getLogger(__name__)


class Clight(object):
    """PN order and inverse powers of the speed of light
    (extensions.py:30-60), read by PNAcc at each call."""

    def __init__(self):
        self._pn_order = 0
        self._clight = None

    @property
    def pn_order(self):
        return self._pn_order

    @pn_order.setter
    def pn_order(self, value):
        self._pn_order = int(value)

    @property
    def clight(self):
        return self._clight

    @clight.setter
    def clight(self, value):
        self._clight = float(value)
        self.inv1 = 1.0 / self._clight
        self.inv2 = self.inv1**2
        self.inv3 = self.inv1**3
        self.inv4 = self.inv1**4
        self.inv5 = self.inv1**5
        self.inv6 = self.inv1**6
        self.inv7 = self.inv1**7


def get_kernel(name, backend, prec, **kwargs):
    """Kernel object for ``name`` (extensions.py:63-76)."""
    if isinstance(backend, str):
        if backend == "HIP":
            from .hip_backend import HIPKernel as Kernel
        else:
            msg = "Inappropriate 'backend': {}. Supported values: ['HIP']"
            raise ValueError(msg.format(backend))
    elif callable(backend):
        Kernel = backend
    else:
        raise ValueError("Inappropriate 'backend': %r" % (backend,))
    logger.debug("Using '%s' from %s precision %s extension module.",
                 name, prec, backend)
    return Kernel(prec, name, **kwargs)


# Interaction counter: None (off) or a Counter of n_i * n_j per extension
# class, for throughput accounting (e.g. the pairs of one block step).
pair_counts = None


def count_pairs(enable=True):
    """Start (reset) or stop counting n_i * n_j per kernel call; returns
    the counter."""
    global pair_counts
    pair_counts = collections.Counter() if enable else None
    return pair_counts


def prepare_args(args, argtypes):
    return [argtype(arg) for (arg, argtype) in zip(args, argtypes)]


def _register(ps, *names):
    for name in names:
        if name not in ps.__dict__:
            ps.register_auxiliary_attribute(name, "real")


class AbstractExtension(object):
    def set_args(self, ips, jps, *args):
        raise NotImplementedError

    def run(self):
        self.kernel.run()

    def get_result(self):
        return self.kernel.map_buffers(self._outargs, self.outargs)

    def calc(self, ips, jps, *args):
        self.set_args(ips, jps, *args)
        self.run()
        return self.get_result()

    def _bind(self, ni, nj, inargs, outargs):
        if pair_counts is not None:
            pair_counts[type(self).__name__] += ni * nj
        self._inargs = inargs
        self._outargs = outargs
        self.inargs = prepare_args(self._inargs, self.argtypes)
        self.outargs = prepare_args(self._outargs, self.restypes)
        self.kernel.set_args(self.inargs + self.outargs)


class Phi(AbstractExtension):
    """Gravitational potential phi_i (extensions.py:100-159)."""

    def __init__(self, backend, prec, **kw):
        self.kernel = get_kernel("phi_kernel", backend, prec, **kw)
        cty = self.kernel.cty
        self.argtypes = ((cty.c_uint,) + (cty.c_real_p,) * 5
                         + (cty.c_uint,) + (cty.c_real_p,) * 5)
        self.restypes = (cty.c_real_p,)

    def set_args(self, ips, jps):
        ni, nj = ips.n, jps.n
        self.kernel.set_gsize(ni, nj)
        _register(ips, "phi")
        self._bind(ni, nj,
                   (ni, ips.mass, ips.rx, ips.ry, ips.rz, ips.eps2,
                    nj, jps.mass, jps.rx, jps.ry, jps.rz, jps.eps2),
                   (ips.phi,))


class Acc(AbstractExtension):
    """Gravitational acceleration (extensions.py:162-233)."""

    def __init__(self, backend, prec, **kw):
        self.kernel = get_kernel("acc_kernel", backend, prec, **kw)
        cty = self.kernel.cty
        self.argtypes = ((cty.c_uint,) + (cty.c_real_p,) * 5
                         + (cty.c_uint,) + (cty.c_real_p,) * 5)
        self.restypes = (cty.c_real_p,) * 3

    def set_args(self, ips, jps):
        ni, nj = ips.n, jps.n
        self.kernel.set_gsize(ni, nj)
        _register(ips, "ax", "ay", "az")
        self._bind(ni, nj,
                   (ni, ips.mass, ips.rx, ips.ry, ips.rz, ips.eps2,
                    nj, jps.mass, jps.rx, jps.ry, jps.rz, jps.eps2),
                   (ips.ax, ips.ay, ips.az))


class AccJerk(AbstractExtension):
    """Acceleration and jerk (extensions.py:236-285)."""

    def __init__(self, backend, prec, **kw):
        self.kernel = get_kernel("acc_jerk_kernel", backend, prec, **kw)
        cty = self.kernel.cty
        self.argtypes = ((cty.c_uint,) + (cty.c_real_p,) * 8
                         + (cty.c_uint,) + (cty.c_real_p,) * 8)
        self.restypes = (cty.c_real_p,) * 6

    def set_args(self, ips, jps):
        ni, nj = ips.n, jps.n
        self.kernel.set_gsize(ni, nj)
        _register(ips, "ax", "ay", "az", "jx", "jy", "jz")
        self._bind(ni, nj,
                   (ni, ips.mass, ips.rx, ips.ry, ips.rz,
                    ips.eps2, ips.vx, ips.vy, ips.vz,
                    nj, jps.mass, jps.rx, jps.ry, jps.rz,
                    jps.eps2, jps.vx, jps.vy, jps.vz),
                   (ips.ax, ips.ay, ips.az, ips.jx, ips.jy, ips.jz))


class SnapCrackle(AbstractExtension):
    """Snap and crackle (extensions.py:288-343)."""

    def __init__(self, backend, prec, **kw):
        self.kernel = get_kernel("snap_crackle_kernel", backend, prec, **kw)
        cty = self.kernel.cty
        self.argtypes = ((cty.c_uint,) + (cty.c_real_p,) * 14
                         + (cty.c_uint,) + (cty.c_real_p,) * 14)
        self.restypes = (cty.c_real_p,) * 6

    def set_args(self, ips, jps):
        ni, nj = ips.n, jps.n
        self.kernel.set_gsize(ni, nj)
        _register(ips, "sx", "sy", "sz", "cx", "cy", "cz")
        self._bind(ni, nj,
                   (ni, ips.mass, ips.rx, ips.ry, ips.rz,
                    ips.eps2, ips.vx, ips.vy, ips.vz,
                    ips.ax, ips.ay, ips.az, ips.jx, ips.jy, ips.jz,
                    nj, jps.mass, jps.rx, jps.ry, jps.rz,
                    jps.eps2, jps.vx, jps.vy, jps.vz,
                    jps.ax, jps.ay, jps.az, jps.jx, jps.jy, jps.jz),
                   (ips.sx, ips.sy, ips.sz, ips.cx, ips.cy, ips.cz))


class Tstep(AbstractExtension):
    """Individual time-steps (extensions.py:346-387)."""

    def __init__(self, backend, prec, **kw):
        self.kernel = get_kernel("tstep_kernel", backend, prec, **kw)
        cty = self.kernel.cty
        self.argtypes = ((cty.c_uint,) + (cty.c_real_p,) * 8
                         + (cty.c_uint,) + (cty.c_real_p,) * 8
                         + (cty.c_real,))
        self.restypes = (cty.c_real_p,) * 2

    def set_args(self, ips, jps, eta):
        ni, nj = ips.n, jps.n
        self.kernel.set_gsize(ni, nj)
        _register(ips, "tstep", "tstepij")
        self._bind(ni, nj,
                   (ni, ips.mass, ips.rx, ips.ry, ips.rz,
                    ips.eps2, ips.vx, ips.vy, ips.vz,
                    nj, jps.mass, jps.rx, jps.ry, jps.rz,
                    jps.eps2, jps.vx, jps.vy, jps.vz,
                    eta),
                   (ips.tstep, ips.tstepij))


class Sakura(AbstractExtension):
    """Sakura pair drift (extensions.py:442-507)."""

    def __init__(self, backend, prec, **kw):
        self.kernel = get_kernel("sakura_kernel", backend, prec, **kw)
        cty = self.kernel.cty
        self.argtypes = ((cty.c_uint,) + (cty.c_real_p,) * 8
                         + (cty.c_uint,) + (cty.c_real_p,) * 8
                         + (cty.c_real, cty.c_int))
        self.restypes = (cty.c_real_p,) * 6

    def set_args(self, ips, jps, dt, flag):
        ni, nj = ips.n, jps.n
        self.kernel.set_gsize(ni, nj)
        _register(ips, "drx", "dry", "drz", "dvx", "dvy", "dvz")
        self._bind(ni, nj,
                   (ni, ips.mass, ips.rx, ips.ry, ips.rz,
                    ips.eps2, ips.vx, ips.vy, ips.vz,
                    nj, jps.mass, jps.rx, jps.ry, jps.rz,
                    jps.eps2, jps.vx, jps.vy, jps.vz,
                    dt, flag),
                   (ips.drx, ips.dry, ips.drz, ips.dvx, ips.dvy, ips.dvz))


class Kepler(AbstractExtension):
    """Exact two-body advance of particles 0 and 1 of ``ips``, in place
    (extensions.py:618-651)."""

    def __init__(self, backend, prec, **kw):
        self.kernel = get_kernel("kepler_solver_kernel", backend, prec, **kw)
        cty = self.kernel.cty
        self.argtypes = (cty.c_real_p,) * 8 + (cty.c_real,)
        self.restypes = (cty.c_real_p,) * 6

    def set_args(self, ips, jps, dt):
        self._bind(ips.n, jps.n,
                   (ips.mass, ips.rx, ips.ry, ips.rz,
                    ips.eps2, ips.vx, ips.vy, ips.vz,
                    dt),
                   (ips.rx, ips.ry, ips.rz, ips.vx, ips.vy, ips.vz))


class NREG_X(AbstractExtension):
    """NREG position step: sum_j m_j r_ij(t + dt), the acceleration at the
    drifted separations and m_i sum_j m_j / R (extensions.py:511-565)."""

    def __init__(self, backend, prec, **kw):
        self.kernel = get_kernel("nreg_Xkernel", backend, prec, **kw)
        cty = self.kernel.cty
        self.argtypes = ((cty.c_uint,) + (cty.c_real_p,) * 8
                         + (cty.c_uint,) + (cty.c_real_p,) * 8
                         + (cty.c_real,))
        self.restypes = (cty.c_real_p,) * 7

    def set_args(self, ips, jps, dt):
        ni, nj = ips.n, jps.n
        self.kernel.set_gsize(ni, nj)
        _register(ips, "mrx", "mry", "mrz", "ax", "ay", "az", "u")
        self._bind(ni, nj,
                   (ni, ips.mass, ips.rx, ips.ry, ips.rz,
                    ips.eps2, ips.vx, ips.vy, ips.vz,
                    nj, jps.mass, jps.rx, jps.ry, jps.rz,
                    jps.eps2, jps.vx, jps.vy, jps.vz,
                    dt),
                   (ips.mrx, ips.mry, ips.mrz, ips.ax, ips.ay, ips.az,
                    ips.u))


class NREG_V(AbstractExtension):
    """NREG velocity step: sum_j m_j v_ij(t + dt) and m_i sum_j m_j v_ij^2
    (extensions.py:568-615)."""

    def __init__(self, backend, prec, **kw):
        self.kernel = get_kernel("nreg_Vkernel", backend, prec, **kw)
        cty = self.kernel.cty
        self.argtypes = ((cty.c_uint,) + (cty.c_real_p,) * 7
                         + (cty.c_uint,) + (cty.c_real_p,) * 7
                         + (cty.c_real,))
        self.restypes = (cty.c_real_p,) * 4

    def set_args(self, ips, jps, dt):
        ni, nj = ips.n, jps.n
        self.kernel.set_gsize(ni, nj)
        _register(ips, "mvx", "mvy", "mvz", "mk")
        self._bind(ni, nj,
                   (ni, ips.mass, ips.vx, ips.vy, ips.vz,
                    ips.ax, ips.ay, ips.az,
                    nj, jps.mass, jps.vx, jps.vy, jps.vz,
                    jps.ax, jps.ay, jps.az,
                    dt),
                   (ips.mvx, ips.mvy, ips.mvz, ips.mk))


class PNAcc(AbstractExtension):
    """Post-Newtonian pair accelerations pnax, pnay, pnaz
    (extensions.py:391-436).  The PN order and 1/c^k come from the module's
    ``clight`` at each call, as in the reference."""

    def __init__(self, backend, prec, **kw):
        self.kernel = get_kernel("pnacc_kernel", backend, prec, **kw)
        cty = self.kernel.cty
        self.argtypes = ((cty.c_uint,) + (cty.c_real_p,) * 8
                         + (cty.c_uint,) + (cty.c_real_p,) * 8
                         + (cty.c_uint,) + (cty.c_real,) * 7)
        self.restypes = (cty.c_real_p,) * 3

    def set_args(self, ips, jps):
        if clight.pn_order > 0 and clight.clight is None:
            raise TypeError("PNAcc: 'clight' is not set")
        ni, nj = ips.n, jps.n
        self.kernel.set_gsize(ni, nj)
        _register(ips, "pnax", "pnay", "pnaz")
        inv = [getattr(clight, "inv%d" % k, 0.0) for k in range(1, 8)]
        self._bind(ni, nj,
                   (ni, ips.mass, ips.rx, ips.ry, ips.rz,
                    ips.eps2, ips.vx, ips.vy, ips.vz,
                    nj, jps.mass, jps.rx, jps.ry, jps.rz,
                    jps.eps2, jps.vx, jps.vy, jps.vz,
                    clight.pn_order) + tuple(inv),
                   (ips.pnax, ips.pnay, ips.pnaz))


backend = "HIP"

clight = Clight()
phi = acc = acc_jerk = snap_crackle = tstep = sakura = kepler = None
nreg_x = nreg_v = pnacc = None


def install(backend_=None, prec=None, **kw):
    """(Re)build the module singletons (extensions.py:654-666) for a backend
    and precision.  Called at import with the HIP backend and
    ``ctype.prec``."""
    global backend, phi, acc, acc_jerk, snap_crackle, tstep, sakura, kepler
    global nreg_x, nreg_v, pnacc
    backend = "HIP" if backend_ is None else backend_
    prec = ctype.prec if prec is None else prec
    phi = Phi(backend, prec, **kw)
    acc = Acc(backend, prec, **kw)
    acc_jerk = AccJerk(backend, prec, **kw)
    snap_crackle = SnapCrackle(backend, prec, **kw)
    tstep = Tstep(backend, prec, **kw)
    sakura = Sakura(backend, prec, **kw)
    kepler = Kepler(backend, prec, **kw)
    nreg_x = NREG_X(backend, prec, **kw)
    nreg_v = NREG_V(backend, prec, **kw)
    pnacc = PNAcc(backend, prec, **kw)


install()
