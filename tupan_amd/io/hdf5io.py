"""HDF5 snapshots and worldlines in the reference's layout
(tupan/io/hdf5io.py:22-146).

A file holds ``Snapshot`` (or ``Snapshot_NNNNNN`` for an integer snapshot
number) and ``Worldline`` groups; each holds one group named after the
particle-system class, ``particlesystem``, whose string attribute ``Class``
names that class, and in it one dataset per member type (``bodies``,
``stars``, ``blackholes``, ``sphs``) of the member's records
(``get_state()``: id, mass, eps2, rx, ry, rz, vx, vy, vz, time, nstep, tstep
and the type's own attributes), with its own ``Class`` attribute.  The
datasets are 1-D, extendible and chunked, shuffled and gzip-compressed, and
a second dump into the same group appends records (hdf5io.py:31-50); a
worldline is a growing record of particle states.

The ``Class`` attribute is the text that ``pickle.dumps(cls, protocol=0)``
gives for a class, decoded as UTF-8 (hdf5io.py:19,47,63): ``c<module>
newline <name> newline p0 newline .``.  Files name the reference's classes
(``tupan.particles.allparticles.ParticleSystem``,
``tupan.particles.body.Bodies``, ...), so they read back into the reference
and compare with its files record by record.  Loading never unpickles: the
attribute is parsed as that text and the class picked by name from the
known particle classes, of tupan or of this package; any other content is
an error.
"""
import numpy as np
import torch

from . import h5lib
from ..lib.utils import ctype
from ..particles.system import ATTRS, ParticleSystem, default_device

__all__ = ["HDF5IO", "class_string", "parse_class_string"]

SYSTEM_CLASS = ("tupan.particles.allparticles", "ParticleSystem")
MEMBER_CLASSES = {
    "bodies": ("tupan.particles.body", "Bodies"),
    "stars": ("tupan.particles.star", "Stars"),
    "blackholes": ("tupan.particles.blackhole", "Blackholes"),
    "sphs": ("tupan.particles.sph", "Sphs"),
}
_MEMBER_BY_CLASS = {cls: key for key, (_, cls) in MEMBER_CLASSES.items()}
_KNOWN_MODULES = ("tupan.", "tupan_amd.")


def class_string(module, name):
    """What pickle protocol 0 writes for a reference to class ``name`` of
    ``module`` (GLOBAL, PUT 0, STOP), as text."""
    return "c%s\n%s\np0\n." % (module, name)


def parse_class_string(text):
    """(module, name) of a ``Class`` attribute, read as protocol-0 pickle
    text without unpickling it."""
    if isinstance(text, bytes):
        text = text.decode("utf-8")
    if not text.startswith("c"):
        raise ValueError("Class attribute is not a protocol-0 class "
                         "reference: %r" % (text[:40],))
    parts = text[1:].split("\n")
    if len(parts) < 3:
        raise ValueError("Class attribute is truncated: %r" % (text[:40],))
    module, name = parts[0], parts[1]
    if not module.startswith(_KNOWN_MODULES):
        raise ValueError("Class %s.%s is not a tupan particle class"
                         % (module, name))
    return module, name


def _base_name(base, snap_number):
    if isinstance(snap_number, int):
        base += "_" + str(snap_number).zfill(6)
    return base


def system_from_states(members, device=None):
    """ParticleSystem holding ``members``, a list of (member type, record
    array) in order: the base attributes concatenated in that order, the
    types' own attributes registered as "real" arrays (zeros for members
    without them)."""
    dev = torch.device(device) if device is not None else default_device()
    n = sum(len(st) for _, st in members)
    arrays = {}
    for name, sct in ATTRS:
        parts = [np.asarray(st[name]) if name in st.dtype.names
                 else np.zeros(len(st)) for _, st in members]
        host = (np.concatenate(parts) if parts else np.zeros(0))
        if sct == "real":
            host = host.astype(ctype.REAL)
        else:
            host = host.astype(np.int64)
        arrays[name] = torch.as_tensor(host, device=dev)
    extras = []
    for _, st in members:
        for name in st.dtype.names:
            if name not in dict(ATTRS) and name not in extras:
                extras.append(name)
    for name in extras:
        parts = [np.asarray(st[name], dtype=ctype.REAL)
                 if name in st.dtype.names else np.zeros(len(st), ctype.REAL)
                 for _, st in members]
        arrays[name] = torch.as_tensor(np.concatenate(parts), device=dev)
    counts = [(key, len(st)) for key, st in members] or [("bodies", 0)]
    ps = ParticleSystem(device=dev, _arrays=arrays, _members=counts)
    assert ps.n == n
    return ps


class HDF5IO(object):
    """HDF5IO(fname, fmode) with the reference's methods; ``fmode`` is an
    h5py file mode ('r', 'r+', 'w', 'w-', 'x', 'a')."""

    def __init__(self, fname, fmode):
        self.fname = fname
        self.fmode = fmode

    def store_dset(self, group, key, obj):
        """Append member ``obj``'s records to dataset ``key`` of ``group``,
        creating it (empty, extendible) first (hdf5io.py:31-50)."""
        state = obj.get_state()
        if key in group:
            dset = group[key]
        else:
            dset = group.create_extendible(key, state.dtype, 0)
        with dset:
            olen = len(dset)
            dset.attrs["Class"] = class_string(*MEMBER_CLASSES[key])
            dset.resize(olen + len(state))
            dset.write(olen, state)

    def _dumpper(self, fobj, ps, base_name):
        with fobj.require_group(base_name) as base_group:
            group_name = type(ps).__name__.lower()
            with base_group.require_group(group_name) as group:
                group.attrs["Class"] = class_string(*SYSTEM_CLASS)
                for (key, obj) in ps.members.items():
                    if obj.n:
                        self.store_dset(group, key, obj)

    def snapshot_dumpper(self, fobj, ps, snap_number):
        """hdf5io.py:52-66."""
        self._dumpper(fobj, ps, _base_name("Snapshot", snap_number))

    def dump_snapshot(self, ps, snap_number=None):
        with h5lib.File(self.fname, self.fmode) as fobj:
            self.snapshot_dumpper(fobj, ps, snap_number)

    def _loader(self, fobj, base_name, device):
        with fobj[base_name] as base:
            names = base.keys()
            if not names:
                raise ValueError("group %r is empty" % base_name)
            with base[names[0]] as group:
                module, name = parse_class_string(group.attrs["Class"])
                if name != SYSTEM_CLASS[1]:
                    raise ValueError("unexpected system class %s.%s"
                                     % (module, name))
                members = []
                for key in group.keys():
                    with group[key] as dset:
                        module, name = parse_class_string(dset.attrs["Class"])
                        if name not in _MEMBER_BY_CLASS:
                            raise ValueError("unexpected member class %s.%s"
                                             % (module, name))
                        members.append((_MEMBER_BY_CLASS[name], dset.read()))
        return system_from_states(members, device)

    def snapshot_loader(self, fobj, snap_number=None, device=None):
        """hdf5io.py:74-87: the first group of the snapshot, its members in
        name order."""
        return self._loader(fobj, _base_name("Snapshot", snap_number),
                            device)

    def load_snapshot(self, snap_number=None, device=None):
        with h5lib.File(self.fname, self.fmode) as fobj:
            return self.snapshot_loader(fobj, snap_number, device)

    def worldline_dumpper(self, fobj, wl):
        """hdf5io.py:97-110."""
        self._dumpper(fobj, wl, "Worldline")

    def dump_worldline(self, wl):
        with h5lib.File(self.fname, self.fmode) as fobj:
            self.worldline_dumpper(fobj, wl)

    def worldline_loader(self, fobj, wl_number=None, device=None):
        """Every record of the worldline, in the order written (the
        reference leaves this unimplemented, hdf5io.py:119-123)."""
        return self._loader(fobj, "Worldline", device)

    def load_worldline(self, wl_number=None, device=None):
        with h5lib.File(self.fname, self.fmode) as fobj:
            return self.worldline_loader(fobj, wl_number, device)

    def to_psdf(self):
        raise NotImplementedError("PSDF output is not provided")
