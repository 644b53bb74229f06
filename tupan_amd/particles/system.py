"""Device-resident particle system (tupan/particles/allparticles.py:24-206 and
tupan/particles/body.py:19-461).

Every attribute is a contiguous 1-D array, one value per particle (SoA), held
as a torch tensor on a HIP device, so the kernels read and write it in HBM
without host round trips.  The interface follows the reference:

  * attributes ``id mass eps2 rx ry rz vx vy vz time nstep tstep``
    (body.py:26-39); ``eps2`` holds half the squared softening length
    (ics/plummer.py:62-63), so a pair sees e2_i + e2_j;
  * ``register_auxiliary_attribute(name, sctype)`` adds a zero-filled array
    (allparticles.py:64-76); kernel outputs are registered this way;
  * ``copy()``, ``append()``, ``ps[mask]``, ``ps[i]``, ``ps[a:b]``,
    ``ps[...] = other`` (allparticles.py:97-206);
  * member types (``ParticleSystem(nbodies, nstars, nbhs, nsphs)``) are
    contiguous ranges of the same arrays, reachable as ``ps.members[name]``
    and ``ps.<name>`` views (allparticles.py:29-62);
  * energies, momenta, centre of mass and the rescalings of body.py:65-461.

Integer attributes (``id``, ``nstep``) are stored as int64 (torch has no
unsigned 64-bit arithmetic on the device).  Reductions that feed energy
diagnostics are summed on the host with ``math.fsum`` (exactly rounded).

``t_curr`` is a class attribute, as in the reference (integrator/__init__.py:
19).
"""
import math

import numpy as np
import torch

from ..lib.utils import ctype

ATTRS = (("id", "uint"), ("mass", "real"), ("eps2", "real"),
         ("rx", "real"), ("ry", "real"), ("rz", "real"),
         ("vx", "real"), ("vy", "real"), ("vz", "real"),
         ("time", "real"), ("nstep", "uint"), ("tstep", "real"))
MEMBER_TYPES = ("bodies", "stars", "blackholes", "sphs")
# Attributes each member type adds to the base ones (star.py:21-28,
# blackhole.py:21-26, sph.py:21-23); all "real".  They exist on a system
# only when registered (or loaded from a file that holds them).
MEMBER_EXTRAS = {
    "bodies": (),
    "stars": ("spinx", "spiny", "spinz", "radius", "age", "metallicity"),
    "blackholes": ("spinx", "spiny", "spinz", "radius"),
    "sphs": ("density",),
}


def default_device():
    if torch.cuda.is_available():
        return torch.device("cuda", torch.cuda.current_device())
    return torch.device("cpu")


def torch_dtype(sctype):
    if sctype == "real":
        return torch.float64 if ctype.REAL == np.float64 else torch.float32
    return torch.int64 if ctype.INT == np.int64 else torch.int32


def fsum(t):
    """Exactly rounded sum of a tensor's values (host math.fsum)."""
    return math.fsum(t.detach().to("cpu", torch.float64).numpy().tolist())


class Member(object):
    """View of one member type: attribute access returns the slice of the
    system's arrays that belongs to it."""

    def __init__(self, system, start, stop, name="bodies"):
        self.__dict__["_ps"] = system
        self.__dict__["_sl"] = slice(start, stop)
        self.__dict__["name"] = name

    @property
    def n(self):
        return self._sl.stop - self._sl.start

    def __len__(self):
        return self.n

    def __getattr__(self, name):
        ps = self.__dict__["_ps"]
        if name in ps.__dict__ and isinstance(ps.__dict__[name], torch.Tensor):
            return ps.__dict__[name][self._sl]
        raise AttributeError(name)

    def __setattr__(self, name, value):
        getattr(self, name)[...] = value

    def get_state(self):
        """Structured array of this member's particles in its type's record
        layout: the base attributes, then the type's own (body.py:664-668
        with the attrs lists of star.py, blackhole.py, sph.py); an own
        attribute the system does not hold is written as zeros."""
        dt = ([(name, ctype.ctypedict[sct]) for name, sct in ATTRS]
              + [(name, ctype.REAL) for name in MEMBER_EXTRAS[self.name]])
        out = np.zeros(self.n, dtype=dt)
        arrays = self.__dict__["_ps"].__dict__
        for name in out.dtype.names:
            if name in arrays and isinstance(arrays[name], torch.Tensor):
                out[name] = arrays[name][self._sl].detach().cpu().numpy()
        return out


class ParticleSystem(object):
    """Particle system with device-resident SoA arrays."""

    t_curr = 0.0
    include_pn_corrections = False

    def __init__(self, nbodies=0, nstars=0, nbhs=0, nsphs=0, device=None,
                 _arrays=None, _members=None):
        dev = torch.device(device) if device is not None else default_device()
        self.__dict__["device"] = dev
        if _arrays is not None:
            self.__dict__.update(_arrays)
            n = int(next(iter(_arrays.values())).shape[0]) if _arrays else 0
            self.__dict__["_member_counts"] = (
                list(_members) if _members is not None else [("bodies", n)])
        else:
            counts = [(name, int(k)) for name, k in
                      zip(MEMBER_TYPES, (nbodies, nstars, nbhs, nsphs)) if k]
            n = sum(k for _, k in counts)
            self.__dict__["_member_counts"] = counts or [("bodies", 0)]
            for name, sct in ATTRS:
                if name == "id":
                    t = torch.arange(n, dtype=torch_dtype(sct), device=dev)
                else:
                    t = torch.zeros(n, dtype=torch_dtype(sct), device=dev)
                self.__dict__[name] = t
        self.__dict__["n"] = n

    # -- structure ------------------------------------------------------------
    @property
    def members(self):
        out = {}
        start = 0
        for name, k in self._member_counts:
            if k:
                out[name] = Member(self, start, start + k, name)
            start += k
        return out

    def __getattr__(self, name):
        if name in MEMBER_TYPES:
            m = self.members
            if name in m:
                return m[name]
        raise AttributeError(name)

    def arrays(self):
        return {k: v for k, v in self.__dict__.items()
                if isinstance(v, torch.Tensor)}

    def __len__(self):
        return self.n

    def __repr__(self):
        return "ParticleSystem(n=%d, device=%s)" % (self.n, self.device)

    def register_auxiliary_attribute(self, attr, sctype):
        """Zero-filled array; "real" attributes take the dtype of the
        system's masses, so a float32 system gets float32 outputs."""
        if attr in self.__dict__:
            raise ValueError("'{0}' is already a registered "
                             "attribute.".format(attr))
        dt = torch_dtype(sctype)
        if sctype == "real" and "mass" in self.__dict__:
            dt = self.__dict__["mass"].dtype
        self.__dict__[attr] = torch.zeros(self.n, dtype=dt,
                                          device=self.device)

    def _new(self, arrays, members=None):
        return type(self)(device=self.device, _arrays=arrays,
                          _members=members)

    def copy(self):
        return self._new({k: v.clone() for k, v in self.arrays().items()},
                         list(self._member_counts))

    def append(self, obj):
        if not obj.n:
            return
        mine = self.arrays()
        theirs = obj.arrays()
        merged = {}
        for k in set(mine) | set(theirs):
            a = mine.get(k)
            b = theirs.get(k)
            ref = a if a is not None else b
            if a is None:
                a = torch.zeros(self.n, dtype=ref.dtype, device=self.device)
            if b is None:
                b = torch.zeros(obj.n, dtype=ref.dtype, device=self.device)
            merged[k] = torch.cat((a, b.to(self.device)))
        counts = list(self._member_counts)
        for name, k in obj._member_counts:
            counts.append((name, k))
        for k in list(self.__dict__):
            if isinstance(self.__dict__[k], torch.Tensor):
                del self.__dict__[k]
        self.__dict__.update(merged)
        self.__dict__["n"] = self.n + obj.n
        self.__dict__["_member_counts"] = _merge_counts(counts)

    def _index(self, slc):
        if isinstance(slc, np.ndarray):
            slc = torch.as_tensor(slc, device=self.device)
        if isinstance(slc, (int, np.integer)):
            i = int(slc)
            if abs(i) > self.n - 1:
                raise IndexError("index {0} out of bounds 0<=index<{1}"
                                 .format(i, self.n))
            return slice(i % self.n, i % self.n + 1)
        if isinstance(slc, torch.Tensor) and slc.device != self.device:
            slc = slc.to(self.device)
        return slc

    def __getitem__(self, slc):
        idx = self._index(slc)
        arrays = {k: v[idx] for k, v in self.arrays().items()}
        if isinstance(idx, slice):
            # slices are views, as numpy slices are in the reference
            return self._new(arrays, [("bodies", int(arrays["id"].shape[0]))])
        arrays = {k: v.contiguous() for k, v in arrays.items()}
        return self._new(arrays, [("bodies", int(arrays["id"].shape[0]))])

    def __setitem__(self, slc, values):
        idx = self._index(slc)
        src = values.arrays()
        for k, v in self.arrays().items():
            if k in src:
                v[idx] = src[k].to(self.device)

    # -- host transfer ----------------------------------------------------------
    def to_numpy(self):
        return {k: v.detach().cpu().numpy() for k, v in self.arrays().items()}

    @classmethod
    def from_numpy(cls, arrays, device=None):
        dev = torch.device(device) if device is not None else default_device()
        ts = {}
        for k, v in arrays.items():
            v = np.asarray(v)
            if v.dtype.kind == "f":
                ts[k] = torch.as_tensor(v.astype(ctype.REAL), device=dev)
            else:
                ts[k] = torch.as_tensor(v.astype(np.int64), device=dev)
        n = len(next(iter(arrays.values())))
        for name, sct in ATTRS:
            if name not in ts:
                ts[name] = (torch.arange(n, dtype=torch_dtype(sct), device=dev)
                            if name == "id" else
                            torch.zeros(n, dtype=torch_dtype(sct), device=dev))
        return cls(device=dev, _arrays=ts, _members=[("bodies", n)])

    def get_state(self):
        """Structured array of the base attributes (body.py:664-668)."""
        dt = [(name, ctype.ctypedict[sct]) for name, sct in ATTRS]
        out = np.zeros(self.n, dtype=dt)
        for name, _ in ATTRS:
            out[name] = self.__dict__[name].detach().cpu().numpy()
        return out

    def set_state(self, array):
        for name in array.dtype.names:
            if name in self.__dict__:
                self.__dict__[name].copy_(torch.as_tensor(
                    np.ascontiguousarray(array[name]),
                    dtype=self.__dict__[name].dtype))

    # -- gravity (body.py:324-361) --------------------------------------------
    def set_tstep(self, ps, eta):
        from ..lib import extensions
        extensions.tstep.calc(self, ps, eta)

    def set_phi(self, ps):
        from ..lib import extensions
        extensions.phi.calc(self, ps)

    def set_acc(self, ps):
        from ..lib import extensions
        extensions.acc.calc(self, ps)

    def set_acc_jerk(self, ps):
        from ..lib import extensions
        extensions.acc_jerk.calc(self, ps)

    def set_snap_crackle(self, ps):
        from ..lib import extensions
        extensions.snap_crackle.calc(self, ps)

    def set_pnacc(self, ps):
        from ..lib import extensions
        extensions.pnacc.calc(self, ps)

    # -- post-Newtonian bookkeeping (body.py:464-528) -------------------------
    # The PN kicks change v with forces the Newtonian diagnostics do not see;
    # pn_ke, pn_mv and pn_am accumulate minus the work, impulse and torque of
    # those forces and pn_mr the drift of pn_mv, and the diagnostics below add
    # them back when include_pn_corrections is set.
    PN_AUX = ("pnax", "pnay", "pnaz", "pn_ke", "pn_mrx", "pn_mry", "pn_mrz",
              "pn_mvx", "pn_mvy", "pn_mvz", "pn_amx", "pn_amy", "pn_amz")

    def _aux(self, *names):
        for name in names:
            if name not in self.__dict__:
                self.register_auxiliary_attribute(name, "real")
        return [self.__dict__[name] for name in names]

    def _pn_force(self):
        return (self.mass * self.pnax, self.mass * self.pnay,
                self.mass * self.pnaz)

    def pn_kick_ke(self, tau):
        """body.py:471-480."""
        ke, = self._aux("pn_ke")
        fx, fy, fz = self._pn_force()
        ke -= (self.vx * fx + self.vy * fy + self.vz * fz) * tau

    def pn_drift_com_r(self, tau):
        """body.py:482-494."""
        mrx, mry, mrz, mvx, mvy, mvz = self._aux(
            "pn_mrx", "pn_mry", "pn_mrz", "pn_mvx", "pn_mvy", "pn_mvz")
        mrx += mvx * tau
        mry += mvy * tau
        mrz += mvz * tau

    def pn_kick_lmom(self, tau):
        """body.py:496-511."""
        mvx, mvy, mvz = self._aux("pn_mvx", "pn_mvy", "pn_mvz")
        fx, fy, fz = self._pn_force()
        mvx -= fx * tau
        mvy -= fy * tau
        mvz -= fz * tau

    def pn_kick_amom(self, tau):
        """body.py:513-528."""
        amx, amy, amz = self._aux("pn_amx", "pn_amy", "pn_amz")
        fx, fy, fz = self._pn_force()
        amx -= (self.ry * fz - self.rz * fy) * tau
        amy -= (self.rz * fx - self.rx * fz) * tau
        amz -= (self.rx * fy - self.ry * fx) * tau

    def min_tstep(self):
        return float(self.tstep.abs().min())

    def max_tstep(self):
        return float(self.tstep.abs().max())

    # -- mass, centre of mass, momenta (body.py:65-244) -------------------------
    @property
    def px(self):
        return self.mass * self.vx

    @property
    def py(self):
        return self.mass * self.vy

    @property
    def pz(self):
        return self.mass * self.vz

    @property
    def total_mass(self):
        return fsum(self.mass)

    def _pn_on(self):
        return bool(self.include_pn_corrections)

    @property
    def com_r(self):
        """body.py:86-107 (PN corrections included when enabled)."""
        mt = self.total_mass
        m = [self.mass * self.rx, self.mass * self.ry, self.mass * self.rz]
        if self._pn_on():
            m = [a + b for a, b in zip(m, self._aux("pn_mrx", "pn_mry",
                                                    "pn_mrz"))]
        return np.array([fsum(x) for x in m]) / mt

    def _lm(self):
        p = [self.px, self.py, self.pz]
        if self._pn_on():
            p = [a + b for a, b in zip(p, self._aux("pn_mvx", "pn_mvy",
                                                    "pn_mvz"))]
        return p

    @property
    def com_v(self):
        """body.py:109-130 (PN corrections included when enabled)."""
        mt = self.total_mass
        return np.array([fsum(x) for x in self._lm()]) / mt

    @property
    def com_linear_momentum(self):
        return self.total_mass * self.com_v

    @property
    def com_angular_momentum(self):
        return self.total_mass * np.cross(self.com_r, self.com_v)

    @property
    def com_kinetic_energy(self):
        return 0.5 * self.total_mass * (self.com_v**2).sum()

    def com_move_to(self, com_r, com_v):
        self.rx += float(com_r[0])
        self.ry += float(com_r[1])
        self.rz += float(com_r[2])
        self.vx += float(com_v[0])
        self.vy += float(com_v[1])
        self.vz += float(com_v[2])

    def com_to_origin(self):
        self.com_move_to(-self.com_r, -self.com_v)

    @property
    def linear_momentum(self):
        """body.py:178-214 (PN corrections included when enabled)."""
        return np.array([fsum(x) for x in self._lm()])

    @property
    def am(self):
        """body.py:216-241 (PN corrections included when enabled)."""
        px, py, pz = self.px, self.py, self.pz
        am = [(self.ry * pz) - (self.rz * py),
              (self.rz * px) - (self.rx * pz),
              (self.rx * py) - (self.ry * px)]
        if self._pn_on():
            am = [a + b for a, b in zip(am, self._aux("pn_amx", "pn_amy",
                                                      "pn_amz"))]
        return torch.stack(am, dim=1)

    @property
    def angular_momentum(self):
        am = self.am
        return np.array([fsum(am[:, 0]), fsum(am[:, 1]), fsum(am[:, 2])])

    # -- energies (body.py:246-321) ---------------------------------------------
    @property
    def ke(self):
        """body.py:259-274 (PN corrections included when enabled)."""
        ke = 0.5 * self.mass * (self.vx**2 + self.vy**2 + self.vz**2)
        if self._pn_on():
            ke = ke + self._aux("pn_ke")[0]
        return ke

    @property
    def kinetic_energy(self):
        return fsum(self.ke)

    @property
    def pe(self):
        self.set_phi(self)
        return self.mass * self.phi

    @property
    def potential_energy(self):
        return 0.5 * fsum(self.pe)

    @property
    def ve(self):
        return 2 * self.ke + self.pe

    @property
    def virial_energy(self):
        return 2 * self.kinetic_energy + self.potential_energy

    # -- length scales and rescaling (body.py:370-461) ------------------------
    @property
    def virial_radius(self):
        mtot = self.total_mass
        pe = self.potential_energy
        return (mtot**2) / (-2 * pe)

    @property
    def radial_size(self):
        c = self.com_r
        rx = self.rx - float(c[0])
        ry = self.ry - float(c[1])
        rz = self.rz - float(c[2])
        inertia = fsum(self.mass * (rx**2 + ry**2 + rz**2))
        return (inertia / self.total_mass)**0.5

    def dynrescale_total_mass(self, total_mass):
        m_ratio = total_mass / self.total_mass
        self.mass *= m_ratio
        self.rx *= m_ratio
        self.ry *= m_ratio
        self.rz *= m_ratio

    def dynrescale_radial_size(self, size):
        r_scale = size / self.radial_size
        v_scale = 1 / r_scale**0.5
        self.rx *= r_scale
        self.ry *= r_scale
        self.rz *= r_scale
        self.vx *= v_scale
        self.vy *= v_scale
        self.vz *= v_scale

    def dynrescale_virial_radius(self, rvir):
        r_scale = rvir / self.virial_radius
        v_scale = 1 / r_scale**0.5
        self.rx *= r_scale
        self.ry *= r_scale
        self.rz *= r_scale
        self.vx *= v_scale
        self.vy *= v_scale
        self.vz *= v_scale

    def scale_to_virial(self):
        ke = self.kinetic_energy
        pe = self.potential_energy
        v_scale = ((-0.5 * pe) / ke)**0.5
        self.vx *= v_scale
        self.vy *= v_scale
        self.vz *= v_scale

    def to_nbody_units(self):
        self.dynrescale_total_mass(1.0)
        self.dynrescale_virial_radius(1.0)


def _merge_counts(counts):
    """Drop empty ranges and join neighbouring ranges of one member type
    (a worldline appends many slices of the same type)."""
    out = []
    for name, k in counts:
        if k == 0:
            continue
        if out and out[-1][0] == name:
            out[-1] = (name, out[-1][1] + k)
        else:
            out.append((name, k))
    return out or [("bodies", 0)]
