"""CPU oracle -- TEST INFRASTRUCTURE ONLY.

ctypes wrappers around the C restatement in ``oracle/oracle.c`` (and, when it
has been built in this checkout, the reference C backend compiled into
``oracle/_ref``).  Only the test suite, ``__graft_entry__.smoke()`` and the
CPU-baseline leg of ``bench.py`` may import this package: it is the checker,
never the thing measured or shipped.

Kernel specifications follow the reference extension layer
(tupan/lib/extensions.py:100-651): the i-side and j-side SoA fields, the
scalar arguments and the outputs of each kernel, in the reference's argument
order (tupan/lib/src/libtupan.h:2-246).
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))

# name -> (i fields, j fields, scalars, outputs)
SPECS = {
    "phi": (("mass", "rx", "ry", "rz", "eps2"),
            ("mass", "rx", "ry", "rz", "eps2"), (), ("phi",)),
    "acc": (("mass", "rx", "ry", "rz", "eps2"),
            ("mass", "rx", "ry", "rz", "eps2"), (), ("ax", "ay", "az")),
    "acc_jerk": (("mass", "rx", "ry", "rz", "eps2", "vx", "vy", "vz"),
                 ("mass", "rx", "ry", "rz", "eps2", "vx", "vy", "vz"), (),
                 ("ax", "ay", "az", "jx", "jy", "jz")),
    "snap_crackle": (("mass", "rx", "ry", "rz", "eps2", "vx", "vy", "vz",
                      "ax", "ay", "az", "jx", "jy", "jz"),
                     ("mass", "rx", "ry", "rz", "eps2", "vx", "vy", "vz",
                      "ax", "ay", "az", "jx", "jy", "jz"), (),
                     ("sx", "sy", "sz", "cx", "cy", "cz")),
    "tstep": (("mass", "rx", "ry", "rz", "eps2", "vx", "vy", "vz"),
              ("mass", "rx", "ry", "rz", "eps2", "vx", "vy", "vz"),
              ("eta",), ("tstep", "tstepij")),
    "nreg_x": (("mass", "rx", "ry", "rz", "eps2", "vx", "vy", "vz"),
               ("mass", "rx", "ry", "rz", "eps2", "vx", "vy", "vz"),
               ("dt",), ("mrx", "mry", "mrz", "ax", "ay", "az", "u")),
    "nreg_v": (("mass", "vx", "vy", "vz", "ax", "ay", "az"),
               ("mass", "vx", "vy", "vz", "ax", "ay", "az"),
               ("dt",), ("mvx", "mvy", "mvz", "mk")),
    "pnacc": (("mass", "rx", "ry", "rz", "eps2", "vx", "vy", "vz"),
              ("mass", "rx", "ry", "rz", "eps2", "vx", "vy", "vz"),
              ("order", "inv1", "inv2", "inv3", "inv4", "inv5", "inv6",
               "inv7"), ("pnax", "pnay", "pnaz")),
    "sakura": (("mass", "rx", "ry", "rz", "eps2", "vx", "vy", "vz"),
               ("mass", "rx", "ry", "rz", "eps2", "vx", "vy", "vz"),
               ("dt", "flag"), ("drx", "dry", "drz", "dvx", "dvy", "dvz")),
}
KEPLER_IN = ("mass", "rx", "ry", "rz", "eps2", "vx", "vy", "vz")
KEPLER_OUT = ("rx", "ry", "rz", "vx", "vy", "vz")


def _field(ps, name):
    if isinstance(ps, dict):
        return ps[name]
    return getattr(ps, name)


class _Lib:
    """One precision of one library (oracle or compiled reference)."""

    def __init__(self, path, prec, symbol):
        if not os.path.exists(path):
            raise FileNotFoundError(
                "%s is not built (run `make -C oracle`)" % path)
        self.lib = ctypes.CDLL(path)
        self.prec = prec
        self.dtype = np.float64 if prec == "float64" else np.float32
        self.c_real = ctypes.c_double if prec == "float64" else ctypes.c_float
        self.c_uint = ctypes.c_ulong if prec == "float64" else ctypes.c_uint
        self.c_int = ctypes.c_long if prec == "float64" else ctypes.c_int
        self._symbol = symbol
        self._fns = {}

    def _fn(self, name):
        if name not in self._fns:
            fn = getattr(self.lib, self._symbol(name))
            fn.restype = None
            self._fns[name] = fn
        return self._fns[name]

    def _arr(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        return a, a.ctypes.data_as(ctypes.c_void_p)

    def pairwise(self, name, ips, jps, *scalars):
        ifl, jfl, sfl, ofl = SPECS[name]
        keep = []
        iargs = []
        for f in ifl:
            a, p = self._arr(_field(ips, f))
            keep.append(a)
            iargs.append(p)
        ni = len(keep[0])
        jargs = []
        for f in jfl:
            a, p = self._arr(_field(jps, f))
            keep.append(a)
            jargs.append(p)
        nj = len(keep[-1])
        sargs = []
        for f, v in zip(sfl, scalars):
            if f == "flag":
                sargs.append(self.c_int(int(v)))
            elif f == "order":
                sargs.append(self.c_uint(int(v)))
            else:
                sargs.append(self.c_real(v))
        outs = [np.zeros(ni, self.dtype) for _ in ofl]
        oargs = [o.ctypes.data_as(ctypes.c_void_p) for o in outs]
        self._fn(name)(self.c_uint(ni), *iargs, self.c_uint(nj), *jargs,
                       *sargs, *oargs)
        return tuple(outs)

    def kepler(self, ps, dt):
        """Two-body advance of particles 0 and 1; returns new (rx..vz)."""
        keep = []
        args = []
        for f in KEPLER_IN:
            a, p = self._arr(_field(ps, f))
            if len(a) < 2:
                raise ValueError("kepler needs two particles")
            keep.append(a)
            args.append(p)
        outs = [np.array(self._arr(_field(ps, f))[0], copy=True)
                for f in KEPLER_OUT]
        oargs = [o.ctypes.data_as(ctypes.c_void_p) for o in outs]
        self._fn("kepler")(*args, self.c_real(dt), *oargs)
        return tuple(outs)

    def __getattr__(self, name):
        if name in SPECS:
            return lambda ips, jps, *s: self.pairwise(name, ips, jps, *s)
        raise AttributeError(name)


class Oracle(_Lib):
    """The C restatement (oracle.c).  ``omp=True`` loads the OpenMP build."""

    def __init__(self, prec="float64", omp=False):
        suffix = "f64" if prec == "float64" else "f32"
        stem = "liboracle_omp_" if omp else "liboracle_"
        path = os.path.join(HERE, stem + suffix + ".so")
        pre = "oracle_omp_" if omp else "oracle_"
        tail = "" if prec == "float64" else "_f32"
        super().__init__(path, prec, lambda n: pre + n + tail)
        self.c_uint = ctypes.c_size_t     # the oracle takes size_t counts
        self.omp = omp

    def counters_reset(self):
        fn = getattr(self.lib, "oracle_counters_reset" +
                     ("" if self.prec == "float64" else "_f32"))
        fn()

    def counters(self):
        """(kepler solves, laguerre iterations, retries, leapfrogs, fails)."""
        out = (ctypes.c_uint64 * 5)()
        fn = getattr(self.lib, "oracle_counters_get" +
                     ("" if self.prec == "float64" else "_f32"))
        fn(out)
        return tuple(int(x) for x in out)

    def threads(self):
        if not self.omp:
            return 1
        fn = getattr(self.lib, "oracle_omp_threads" +
                     ("" if self.prec == "float64" else "_f32"))
        fn.restype = ctypes.c_int
        return int(fn())

    def kepler_rel(self, dt, m, e2, p0, v0):
        """Advance one relative orbit (p0, v0) by dt; returns (p1, v1)."""
        fn = getattr(self.lib, "oracle_kepler_rel" +
                     ("" if self.prec == "float64" else "_f32"))
        fn.restype = None
        p0, pp = self._arr(p0)
        v0, vp = self._arr(v0)
        p1 = np.zeros(3, self.dtype)
        v1 = np.zeros(3, self.dtype)
        fn(self.c_real(dt), self.c_real(m), self.c_real(e2), pp, vp,
           p1.ctypes.data_as(ctypes.c_void_p),
           v1.ctypes.data_as(ctypes.c_void_p))
        return p1, v1

    def laguerre_from(self, s0, dt, r, rv, m, alpha):
        """Laguerre solve from guess s0: (err, s, iterations)."""
        fn = getattr(self.lib, "oracle_laguerre_from" +
                     ("" if self.prec == "float64" else "_f32"))
        fn.restype = self.c_int
        s = self.c_real()
        it = self.c_int()
        err = fn(self.c_real(s0), self.c_real(dt), self.c_real(r),
                 self.c_real(rv), self.c_real(m), self.c_real(alpha),
                 ctypes.byref(s), ctypes.byref(it))
        return int(err), float(s.value), int(it.value)


class Reference(_Lib):
    """The reference C backend compiled from its own sources into _ref/."""

    def __init__(self, prec="float64"):
        suffix = "f64" if prec == "float64" else "f32"
        path = os.path.join(HERE, "_ref", "libtupan_ref_%s.so" % suffix)
        names = {"phi": "phi_kernel", "acc": "acc_kernel",
                 "acc_jerk": "acc_jerk_kernel",
                 "snap_crackle": "snap_crackle_kernel",
                 "tstep": "tstep_kernel", "sakura": "sakura_kernel",
                 "nreg_x": "nreg_Xkernel", "nreg_v": "nreg_Vkernel",
                 "pnacc": "pnacc_kernel", "kepler": "kepler_solver_kernel"}
        super().__init__(path, prec, lambda n: names[n])


def reference_available(prec="float64"):
    suffix = "f64" if prec == "float64" else "f32"
    return os.path.exists(os.path.join(HERE, "_ref",
                                       "libtupan_ref_%s.so" % suffix))
