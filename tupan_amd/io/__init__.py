"""Snapshot and worldline files (tupan/io/__init__.py:15-97).

``IO(fname, fmode)`` picks the format from the file name, as the reference
does; HDF5 (``.hdf5``, ``hdf5io.HDF5IO``) is the format provided.  PSDF
(``.psdf``) is not: its writer in the reference cannot be constructed
(``PSDFIO(fname)`` takes one argument and is called with two).
"""
import logging
import os
import sys
from warnings import warn

from . import hdf5io  # noqa: F401
from .hdf5io import HDF5IO

__all__ = ["IO", "HDF5IO"]


class IO(object):
    PROVIDED_FORMATS = ["hdf5"]

    def __init__(self, fname, fmode):
        self.fname = fname
        self.fmode = fmode

    def _writer(self):
        if self.fname.endswith(".hdf5"):
            return HDF5IO(self.fname, self.fmode)
        if self.fname.endswith(".psdf"):
            raise NotImplementedError("PSDF files are not provided; "
                                      "choose from: %s"
                                      % (self.PROVIDED_FORMATS,))
        raise NotImplementedError(
            "Unknown format: '{}'. Choose from: {}".format(
                self.fname.rpartition('.')[-1], self.PROVIDED_FORMATS))

    def _load(self, what, *args, **kwargs):
        logger = logging.getLogger(__name__)
        if not os.path.exists(self.fname):
            warn("No such file or directory: '{}'".format(self.fname),
                 stacklevel=3)
            sys.exit()
        try:
            return getattr(HDF5IO(self.fname, self.fmode), what)(
                *args, **kwargs)
        except Exception as exc:
            logger.exception(str(exc))
        raise ValueError("File not in a supported format.")

    def dump_snapshot(self, *args, **kwargs):
        self._writer().dump_snapshot(*args, **kwargs)

    def load_snapshot(self, *args, **kwargs):
        """A missing file warns and exits, and a file no loader reads raises
        ValueError, as in the reference (io/__init__.py:37-55)."""
        return self._load("load_snapshot", *args, **kwargs)

    def dump_worldline(self, *args, **kwargs):
        self._writer().dump_worldline(*args, **kwargs)

    def load_worldline(self, *args, **kwargs):
        return self._load("load_worldline", *args, **kwargs)
