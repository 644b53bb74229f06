"""A small binding of the HDF5 C library (1.10 or later) through ctypes:
files, groups, chunked and filtered 1-D datasets of compound records, and
string attributes -- the subset that the snapshot and worldline layout of
tupan/io/hdf5io.py:31-115 uses, with the storage choices that h5py makes
for it (h5py 3.x, ``h5py/_hl/filters.py``: ``guess_chunk`` for the chunk
shape, gzip level 4 after the shuffle filter, fill time "alloc" for chunked
data; ``attrs[name] = str`` as a scalar variable-length UTF-8 string; link
names created with UTF-8 encoding).

The library is found, in this order, at ``$TUPAN_HDF5_LIB``, by
``ctypes.util.find_library("hdf5")``, or in the lib directory of a conda or
Debian HDF5 install; without it every call raises ``HDF5Unavailable``.
Errors of the library raise ``OSError`` naming the failing call; the
library's own error printing is switched off.
"""
import ctypes
import ctypes.util
import glob
import os
import sys

import numpy as np

hid_t = ctypes.c_int64
herr_t = ctypes.c_int
htri_t = ctypes.c_int
hsize_t = ctypes.c_ulonglong

H5F_ACC_RDONLY = 0x0000
H5F_ACC_RDWR = 0x0001
H5F_ACC_TRUNC = 0x0002
H5F_ACC_EXCL = 0x0004
H5P_DEFAULT = 0
H5S_ALL = 0
H5S_SCALAR = 0
H5S_SELECT_SET = 0
H5S_UNLIMITED = (1 << 64) - 1
H5T_INTEGER = 0
H5T_FLOAT = 1
H5T_STRING = 3
H5T_COMPOUND = 6
H5T_ORDER_LE = 0
H5T_SGN_NONE = 0
H5T_CSET_UTF8 = 1
H5T_VARIABLE = (1 << 64) - 1
H5I_GROUP = 2
H5I_DATASET = 5
H5D_FILL_TIME_ALLOC = 0
H5Z_FILTER_DEFLATE = 1
H5Z_FILTER_SHUFFLE = 2
H5_INDEX_NAME = 0
H5_ITER_INC = 0

# h5py/_hl/filters.py: chunk-size guess constants and the default gzip level
CHUNK_BASE = 16 * 1024
CHUNK_MIN = 8 * 1024
CHUNK_MAX = 1024 * 1024
DEFAULT_GZIP = 4


class HDF5Unavailable(ImportError):
    pass


class _GInfo(ctypes.Structure):
    _fields_ = [("storage_type", ctypes.c_int), ("nlinks", hsize_t),
                ("max_corder", ctypes.c_int64), ("mounted", ctypes.c_bool)]


_SIGS = {
    "H5open": (herr_t, []),
    "H5get_libversion": (herr_t, [ctypes.POINTER(ctypes.c_uint)] * 3),
    "H5Eset_auto2": (herr_t, [hid_t, ctypes.c_void_p, ctypes.c_void_p]),
    "H5free_memory": (herr_t, [ctypes.c_void_p]),
    "H5Fcreate": (hid_t, [ctypes.c_char_p, ctypes.c_uint, hid_t, hid_t]),
    "H5Fopen": (hid_t, [ctypes.c_char_p, ctypes.c_uint, hid_t]),
    "H5Fclose": (herr_t, [hid_t]),
    "H5Gcreate2": (hid_t, [hid_t, ctypes.c_char_p, hid_t, hid_t, hid_t]),
    "H5Gget_info": (herr_t, [hid_t, ctypes.POINTER(_GInfo)]),
    "H5Lexists": (htri_t, [hid_t, ctypes.c_char_p, hid_t]),
    "H5Lget_name_by_idx": (ctypes.c_ssize_t,
                           [hid_t, ctypes.c_char_p, ctypes.c_int,
                            ctypes.c_int, hsize_t, ctypes.c_char_p,
                            ctypes.c_size_t, hid_t]),
    "H5Oopen": (hid_t, [hid_t, ctypes.c_char_p, hid_t]),
    "H5Oclose": (herr_t, [hid_t]),
    "H5Iget_type": (ctypes.c_int, [hid_t]),
    "H5Pcreate": (hid_t, [hid_t]),
    "H5Pclose": (herr_t, [hid_t]),
    "H5Pset_chunk": (herr_t, [hid_t, ctypes.c_int, ctypes.POINTER(hsize_t)]),
    "H5Pget_chunk": (ctypes.c_int, [hid_t, ctypes.c_int,
                                    ctypes.POINTER(hsize_t)]),
    "H5Pset_fill_time": (herr_t, [hid_t, ctypes.c_int]),
    "H5Pset_shuffle": (herr_t, [hid_t]),
    "H5Pset_deflate": (herr_t, [hid_t, ctypes.c_uint]),
    "H5Pset_char_encoding": (herr_t, [hid_t, ctypes.c_int]),
    "H5Pget_nfilters": (ctypes.c_int, [hid_t]),
    "H5Pget_filter2": (ctypes.c_int,
                       [hid_t, ctypes.c_uint, ctypes.POINTER(ctypes.c_uint),
                        ctypes.POINTER(ctypes.c_size_t),
                        ctypes.POINTER(ctypes.c_uint), ctypes.c_size_t,
                        ctypes.c_char_p, ctypes.POINTER(ctypes.c_uint)]),
    "H5Screate": (hid_t, [ctypes.c_int]),
    "H5Screate_simple": (hid_t, [ctypes.c_int, ctypes.POINTER(hsize_t),
                                 ctypes.POINTER(hsize_t)]),
    "H5Sget_simple_extent_ndims": (ctypes.c_int, [hid_t]),
    "H5Sget_simple_extent_dims": (ctypes.c_int,
                                  [hid_t, ctypes.POINTER(hsize_t),
                                   ctypes.POINTER(hsize_t)]),
    "H5Sselect_hyperslab": (herr_t, [hid_t, ctypes.c_int,
                                     ctypes.POINTER(hsize_t),
                                     ctypes.POINTER(hsize_t),
                                     ctypes.POINTER(hsize_t),
                                     ctypes.POINTER(hsize_t)]),
    "H5Sclose": (herr_t, [hid_t]),
    "H5Dcreate2": (hid_t, [hid_t, ctypes.c_char_p, hid_t, hid_t, hid_t,
                           hid_t, hid_t]),
    "H5Dopen2": (hid_t, [hid_t, ctypes.c_char_p, hid_t]),
    "H5Dclose": (herr_t, [hid_t]),
    "H5Dget_space": (hid_t, [hid_t]),
    "H5Dget_type": (hid_t, [hid_t]),
    "H5Dget_create_plist": (hid_t, [hid_t]),
    "H5Dset_extent": (herr_t, [hid_t, ctypes.POINTER(hsize_t)]),
    "H5Dwrite": (herr_t, [hid_t, hid_t, hid_t, hid_t, hid_t,
                          ctypes.c_void_p]),
    "H5Dread": (herr_t, [hid_t, hid_t, hid_t, hid_t, hid_t,
                         ctypes.c_void_p]),
    "H5Tcreate": (hid_t, [ctypes.c_int, ctypes.c_size_t]),
    "H5Tcopy": (hid_t, [hid_t]),
    "H5Tclose": (herr_t, [hid_t]),
    "H5Tinsert": (herr_t, [hid_t, ctypes.c_char_p, ctypes.c_size_t, hid_t]),
    "H5Tset_size": (herr_t, [hid_t, ctypes.c_size_t]),
    "H5Tset_cset": (herr_t, [hid_t, ctypes.c_int]),
    "H5Tget_class": (ctypes.c_int, [hid_t]),
    "H5Tget_size": (ctypes.c_size_t, [hid_t]),
    "H5Tget_order": (ctypes.c_int, [hid_t]),
    "H5Tget_sign": (ctypes.c_int, [hid_t]),
    "H5Tget_cset": (ctypes.c_int, [hid_t]),
    "H5Tis_variable_str": (htri_t, [hid_t]),
    "H5Tget_nmembers": (ctypes.c_int, [hid_t]),
    "H5Tget_member_name": (ctypes.c_void_p, [hid_t, ctypes.c_uint]),
    "H5Tget_member_offset": (ctypes.c_size_t, [hid_t, ctypes.c_uint]),
    "H5Tget_member_type": (hid_t, [hid_t, ctypes.c_uint]),
    "H5Aexists": (htri_t, [hid_t, ctypes.c_char_p]),
    "H5Adelete": (herr_t, [hid_t, ctypes.c_char_p]),
    "H5Acreate2": (hid_t, [hid_t, ctypes.c_char_p, hid_t, hid_t, hid_t,
                           hid_t]),
    "H5Aopen": (hid_t, [hid_t, ctypes.c_char_p, hid_t]),
    "H5Aclose": (herr_t, [hid_t]),
    "H5Awrite": (herr_t, [hid_t, hid_t, ctypes.c_void_p]),
    "H5Aread": (herr_t, [hid_t, hid_t, ctypes.c_void_p]),
    "H5Aget_type": (hid_t, [hid_t]),
}

_GLOBALS = ("H5T_IEEE_F64LE_g", "H5T_IEEE_F32LE_g", "H5T_STD_I64LE_g",
            "H5T_STD_U64LE_g", "H5T_STD_I32LE_g", "H5T_STD_U32LE_g",
            "H5T_STD_I8LE_g", "H5T_STD_U8LE_g", "H5T_STD_I16LE_g",
            "H5T_STD_U16LE_g", "H5T_C_S1_g",
            "H5P_CLS_DATASET_CREATE_ID_g", "H5P_CLS_LINK_CREATE_ID_g")


def _candidates():
    env = os.environ.get("TUPAN_HDF5_LIB")
    if env:
        yield env
        return
    found = ctypes.util.find_library("hdf5")
    if found:
        yield found
    prefixes = [sys.prefix, os.environ.get("CONDA_PREFIX", ""), "/opt/conda"]
    for p in prefixes:
        if p:
            for f in sorted(glob.glob(os.path.join(p, "lib", "libhdf5.so*"))):
                yield f
    for pat in ("/usr/lib/x86_64-linux-gnu/hdf5/serial/libhdf5.so*",
                "/usr/lib/x86_64-linux-gnu/libhdf5_serial.so*",
                "/usr/lib64/libhdf5.so*", "/usr/local/lib/libhdf5.so*"):
        for f in sorted(glob.glob(pat)):
            yield f


class _Lib(object):
    def __init__(self):
        tried = []
        lib = None
        for path in _candidates():
            tried.append(path)
            try:
                lib = ctypes.CDLL(path)
                break
            except OSError:
                lib = None
        if lib is None:
            raise HDF5Unavailable(
                "the HDF5 C library was not found (tried: %s); set "
                "TUPAN_HDF5_LIB to its path" % (", ".join(tried) or "none"))
        self.path = path
        for name, (res, args) in _SIGS.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        self.lib = lib
        ver = [ctypes.c_uint() for _ in range(3)]
        lib.H5get_libversion(*[ctypes.byref(v) for v in ver])
        self.version = tuple(v.value for v in ver)
        if self.version < (1, 10, 0):
            raise HDF5Unavailable(
                "HDF5 %d.%d.%d at %s: 1.10 or later is needed (64-bit "
                "identifiers)" % (self.version + (path,)))
        lib.H5open()
        lib.H5Eset_auto2(0, None, None)
        self.g = {k[:-2]: hid_t.in_dll(lib, k).value for k in _GLOBALS}

    def __getattr__(self, name):
        fn = getattr(self.lib, name)

        def call(*args):
            r = fn(*args)
            if isinstance(r, int) and r < 0:
                raise OSError("HDF5 call %s failed (%s)" % (name, self.path))
            return r
        return call


_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        _LIB = _Lib()
    return _LIB


def available():
    try:
        lib()
        return True
    except HDF5Unavailable:
        return False


def guess_chunk(shape, typesize):
    """Chunk shape for a new chunked dataset, as h5py picks it for
    ``chunks=True`` (h5py/_hl/filters.py ``guess_chunk``): an axis of size 0
    (an extendible dataset created empty) counts as 1024; the PyTables
    target size CHUNK_BASE * 2**log10(MiB of the dataset), clamped to
    [CHUNK_MIN, CHUNK_MAX]; the axes halved in turn (rounding up) until the
    chunk is below the target, or within 50 % of it, and below CHUNK_MAX."""
    shape = tuple((x if x != 0 else 1024) for x in shape)
    ndims = len(shape)
    if ndims == 0:
        raise ValueError("Chunks not allowed for scalar datasets.")
    chunks = np.array(shape, dtype="=f8")
    dset_size = np.prod(chunks) * typesize
    target = CHUNK_BASE * (2 ** np.log10(dset_size / (1024.0 * 1024)))
    target = min(max(target, CHUNK_MIN), CHUNK_MAX)
    idx = 0
    while True:
        chunk_bytes = np.prod(chunks) * typesize
        if ((chunk_bytes < target
             or abs(chunk_bytes - target) / target < 0.5)
                and chunk_bytes < CHUNK_MAX):
            break
        if np.prod(chunks) == 1:
            break
        chunks[idx % ndims] = np.ceil(chunks[idx % ndims] / 2.0)
        idx += 1
    return tuple(int(x) for x in chunks)


# -- types -----------------------------------------------------------------
def _scalar_type(dt):
    L = lib()
    dt = np.dtype(dt)
    if dt.byteorder == ">":
        raise TypeError("big-endian fields are not supported: %r" % (dt,))
    key = {("f", 8): "H5T_IEEE_F64LE", ("f", 4): "H5T_IEEE_F32LE",
           ("i", 8): "H5T_STD_I64LE", ("u", 8): "H5T_STD_U64LE",
           ("i", 4): "H5T_STD_I32LE", ("u", 4): "H5T_STD_U32LE",
           ("i", 2): "H5T_STD_I16LE", ("u", 2): "H5T_STD_U16LE",
           ("i", 1): "H5T_STD_I8LE", ("u", 1): "H5T_STD_U8LE"}.get(
               (dt.kind, dt.itemsize))
    if key is None:
        raise TypeError("unsupported field type %r" % (dt,))
    return L.H5Tcopy(L.g[key])


def compound_type(dt):
    """HDF5 compound type of a numpy structured dtype: one member per field
    at the field's offset, record size = itemsize (h5py's mapping)."""
    L = lib()
    dt = np.dtype(dt)
    tid = L.H5Tcreate(H5T_COMPOUND, dt.itemsize)
    for name in dt.names:
        fdt, off = dt.fields[name][:2]
        st = _scalar_type(fdt)
        try:
            L.H5Tinsert(tid, name.encode("utf-8"), off, st)
        finally:
            L.H5Tclose(st)
    return tid


def _member_dtype(tid):
    L = lib()
    cls = L.H5Tget_class(tid)
    size = L.H5Tget_size(tid)
    if L.H5Tget_order(tid) != H5T_ORDER_LE and size > 1:
        raise TypeError("big-endian members are not supported")
    if cls == H5T_FLOAT:
        return np.dtype("<f%d" % size)
    if cls == H5T_INTEGER:
        sign = L.H5Tget_sign(tid)
        return np.dtype(("<u%d" if sign == H5T_SGN_NONE else "<i%d") % size)
    raise TypeError("unsupported member class %d" % cls)


def numpy_dtype(tid):
    """numpy structured dtype of an HDF5 compound type (fields in member
    order, at the members' offsets)."""
    L = lib()
    if L.H5Tget_class(tid) != H5T_COMPOUND:
        return _member_dtype(tid)
    names, formats, offsets = [], [], []
    for k in range(L.H5Tget_nmembers(tid)):
        p = L.H5Tget_member_name(tid, k)
        try:
            names.append(ctypes.string_at(p).decode("utf-8"))
        finally:
            L.H5free_memory(p)
        mt = L.H5Tget_member_type(tid, k)
        try:
            formats.append(_member_dtype(mt))
        finally:
            L.H5Tclose(mt)
        offsets.append(L.H5Tget_member_offset(tid, k))
    return np.dtype({"names": names, "formats": formats,
                     "offsets": offsets, "itemsize": L.H5Tget_size(tid)})


def _dims(*vals):
    return (hsize_t * len(vals))(*vals)


# -- objects -----------------------------------------------------------------
class _Object(object):
    def __init__(self, hid):
        self.id = hid

    @property
    def attrs(self):
        return Attrs(self.id)


class Attrs(object):
    """String attributes of an object (``attrs[name]`` as h5py gives it:
    a str for a UTF-8 string, bytes for an ASCII one)."""

    def __init__(self, oid):
        self.oid = oid

    def __contains__(self, name):
        return lib().H5Aexists(self.oid, name.encode("utf-8")) > 0

    def __setitem__(self, name, value):
        L = lib()
        if not isinstance(value, str):
            raise TypeError("only str attributes are written")
        key = name.encode("utf-8")
        if L.H5Aexists(self.oid, key) > 0:
            L.H5Adelete(self.oid, key)
        tid = L.H5Tcopy(L.g["H5T_C_S1"])
        sid = L.H5Screate(H5S_SCALAR)
        try:
            L.H5Tset_size(tid, H5T_VARIABLE)
            L.H5Tset_cset(tid, H5T_CSET_UTF8)
            aid = L.H5Acreate2(self.oid, key, tid, sid, H5P_DEFAULT,
                               H5P_DEFAULT)
            try:
                buf = ctypes.c_char_p(value.encode("utf-8"))
                L.H5Awrite(aid, tid, ctypes.byref(buf))
            finally:
                L.H5Aclose(aid)
        finally:
            L.H5Sclose(sid)
            L.H5Tclose(tid)

    def __getitem__(self, name):
        L = lib()
        key = name.encode("utf-8")
        if L.H5Aexists(self.oid, key) <= 0:
            raise KeyError(name)
        aid = L.H5Aopen(self.oid, key, H5P_DEFAULT)
        try:
            ftype = L.H5Aget_type(aid)
            try:
                if L.H5Tget_class(ftype) != H5T_STRING:
                    raise TypeError("attribute %r is not a string" % name)
                utf8 = L.H5Tget_cset(ftype) == H5T_CSET_UTF8
                mtype = L.H5Tcopy(ftype)
                try:
                    if L.H5Tis_variable_str(ftype) > 0:
                        p = ctypes.c_void_p()
                        L.H5Aread(aid, mtype, ctypes.byref(p))
                        raw = ctypes.string_at(p.value) if p.value else b""
                        if p.value:
                            L.H5free_memory(p)
                    else:
                        n = L.H5Tget_size(ftype)
                        buf = ctypes.create_string_buffer(n)
                        L.H5Aread(aid, mtype, buf)
                        raw = buf.raw.rstrip(b"\0")
                finally:
                    L.H5Tclose(mtype)
            finally:
                L.H5Tclose(ftype)
        finally:
            L.H5Aclose(aid)
        return raw.decode("utf-8") if utf8 else raw


class Group(_Object):
    def _lcpl(self):
        L = lib()
        lcpl = L.H5Pcreate(L.g["H5P_CLS_LINK_CREATE_ID"])
        L.H5Pset_char_encoding(lcpl, H5T_CSET_UTF8)
        return lcpl

    def __contains__(self, name):
        return lib().H5Lexists(self.id, name.encode("utf-8"),
                               H5P_DEFAULT) > 0

    def keys(self):
        """Link names in increasing name order (h5py's default order)."""
        L = lib()
        info = _GInfo()
        L.H5Gget_info(self.id, ctypes.byref(info))
        out = []
        for k in range(info.nlinks):
            n = L.H5Lget_name_by_idx(self.id, b".", H5_INDEX_NAME,
                                     H5_ITER_INC, k, None, 0, H5P_DEFAULT)
            buf = ctypes.create_string_buffer(n + 1)
            L.H5Lget_name_by_idx(self.id, b".", H5_INDEX_NAME, H5_ITER_INC,
                                 k, buf, n + 1, H5P_DEFAULT)
            out.append(buf.value.decode("utf-8"))
        return out

    def __getitem__(self, name):
        L = lib()
        key = name.encode("utf-8")
        if L.H5Lexists(self.id, key, H5P_DEFAULT) <= 0:
            raise KeyError(name)
        oid = L.H5Oopen(self.id, key, H5P_DEFAULT)
        kind = L.H5Iget_type(oid)
        if kind == H5I_GROUP:
            return Group(oid)
        if kind == H5I_DATASET:
            return Dataset(oid)
        L.H5Oclose(oid)
        raise TypeError("%r is neither a group nor a dataset" % name)

    def items(self):
        return [(k, self[k]) for k in self.keys()]

    def values(self):
        return [self[k] for k in self.keys()]

    def require_group(self, name):
        if name in self:
            g = self[name]
            if not isinstance(g, Group):
                g.close()
                raise TypeError("%r exists and is not a group" % name)
            return g
        L = lib()
        lcpl = self._lcpl()
        try:
            return Group(L.H5Gcreate2(self.id, name.encode("utf-8"), lcpl,
                                      H5P_DEFAULT, H5P_DEFAULT))
        finally:
            L.H5Pclose(lcpl)

    def create_extendible(self, name, dtype, length=0, compression=True):
        """New 1-D dataset of ``length`` records, extendible without limit,
        chunked by ``guess_chunk``; with ``compression`` the shuffle and
        gzip (level 4) filters, in that order."""
        L = lib()
        dtype = np.dtype(dtype)
        chunks = guess_chunk((length,), dtype.itemsize)
        tid = compound_type(dtype)
        sid = L.H5Screate_simple(1, _dims(length), _dims(H5S_UNLIMITED))
        dcpl = L.H5Pcreate(L.g["H5P_CLS_DATASET_CREATE_ID"])
        lcpl = self._lcpl()
        try:
            L.H5Pset_chunk(dcpl, 1, _dims(*chunks))
            L.H5Pset_fill_time(dcpl, H5D_FILL_TIME_ALLOC)
            if compression:
                L.H5Pset_shuffle(dcpl)
                L.H5Pset_deflate(dcpl, DEFAULT_GZIP)
            did = L.H5Dcreate2(self.id, name.encode("utf-8"), tid, sid, lcpl,
                               dcpl, H5P_DEFAULT)
        finally:
            L.H5Pclose(lcpl)
            L.H5Pclose(dcpl)
            L.H5Sclose(sid)
            L.H5Tclose(tid)
        return Dataset(did)

    def close(self):
        if self.id is not None:
            lib().H5Oclose(self.id)
            self.id = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


class Dataset(_Object):
    @property
    def shape(self):
        L = lib()
        sid = L.H5Dget_space(self.id)
        try:
            nd = L.H5Sget_simple_extent_ndims(sid)
            dims = (hsize_t * max(nd, 1))()
            L.H5Sget_simple_extent_dims(sid, dims, None)
            return tuple(int(dims[k]) for k in range(nd))
        finally:
            L.H5Sclose(sid)

    def __len__(self):
        return self.shape[0]

    @property
    def dtype(self):
        L = lib()
        tid = L.H5Dget_type(self.id)
        try:
            return numpy_dtype(tid)
        finally:
            L.H5Tclose(tid)

    @property
    def chunks(self):
        L = lib()
        pid = L.H5Dget_create_plist(self.id)
        try:
            dims = (hsize_t * 8)()
            nd = L.H5Pget_chunk(pid, 8, dims)
            return tuple(int(dims[k]) for k in range(nd))
        except OSError:
            return None
        finally:
            L.H5Pclose(pid)

    @property
    def filters(self):
        """(filter id, client values) in pipeline order."""
        L = lib()
        pid = L.H5Dget_create_plist(self.id)
        out = []
        try:
            for k in range(L.H5Pget_nfilters(pid)):
                flags = ctypes.c_uint()
                n = ctypes.c_size_t(8)
                cd = (ctypes.c_uint * 8)()
                name = ctypes.create_string_buffer(64)
                conf = ctypes.c_uint()
                fid = L.H5Pget_filter2(pid, k, ctypes.byref(flags),
                                       ctypes.byref(n), cd, 64, name,
                                       ctypes.byref(conf))
                out.append((fid, tuple(cd[i] for i in range(n.value))))
        finally:
            L.H5Pclose(pid)
        return out

    def resize(self, n):
        lib().H5Dset_extent(self.id, _dims(n))

    def write(self, start, data):
        """Write records ``data`` (a structured array whose fields are a
        superset of the dataset's, converted by name) at ``start``."""
        L = lib()
        dt = self.dtype
        data = np.asarray(data)
        rec = np.zeros(len(data), dtype=dt)
        for name in dt.names:
            if name in data.dtype.names:
                rec[name] = data[name]
        n = len(rec)
        if n == 0:
            return
        tid = compound_type(dt)
        fsid = L.H5Dget_space(self.id)
        msid = L.H5Screate_simple(1, _dims(n), None)
        try:
            L.H5Sselect_hyperslab(fsid, H5S_SELECT_SET, _dims(start), None,
                                  _dims(n), None)
            L.H5Dwrite(self.id, tid, msid, fsid, H5P_DEFAULT,
                       rec.ctypes.data_as(ctypes.c_void_p))
        finally:
            L.H5Sclose(msid)
            L.H5Sclose(fsid)
            L.H5Tclose(tid)

    def read(self):
        L = lib()
        dt = self.dtype
        out = np.zeros(self.shape[0], dtype=dt)
        if len(out):
            tid = compound_type(dt)
            try:
                L.H5Dread(self.id, tid, H5S_ALL, H5S_ALL, H5P_DEFAULT,
                          out.ctypes.data_as(ctypes.c_void_p))
            finally:
                L.H5Tclose(tid)
        return out

    def __getitem__(self, key):
        if key is Ellipsis or key == slice(None):
            return self.read()
        return self.read()[key]

    def close(self):
        if self.id is not None:
            lib().H5Oclose(self.id)
            self.id = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


class File(Group):
    """An HDF5 file with h5py's modes: 'r' read only, 'r+' read/write (must
    exist), 'w' create or truncate, 'w-' or 'x' create (must not exist),
    'a' read/write, created if missing."""

    def __init__(self, fname, mode="r"):
        L = lib()
        path = os.fsencode(fname)
        if mode == "r":
            fid = L.H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT)
        elif mode == "r+":
            fid = L.H5Fopen(path, H5F_ACC_RDWR, H5P_DEFAULT)
        elif mode == "w":
            fid = L.H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT)
        elif mode in ("w-", "x"):
            fid = L.H5Fcreate(path, H5F_ACC_EXCL, H5P_DEFAULT, H5P_DEFAULT)
        elif mode == "a":
            if os.path.exists(fname):
                fid = L.H5Fopen(path, H5F_ACC_RDWR, H5P_DEFAULT)
            else:
                fid = L.H5Fcreate(path, H5F_ACC_EXCL, H5P_DEFAULT,
                                  H5P_DEFAULT)
        else:
            raise ValueError("Invalid mode; must be one of r, r+, w, w-, x, a")
        self.fid = fid
        Group.__init__(self, L.H5Oopen(fid, b"/", H5P_DEFAULT))

    def close(self):
        Group.close(self)
        if self.fid is not None:
            lib().H5Fclose(self.fid)
            self.fid = None
