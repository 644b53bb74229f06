"""The C ABI: both libraries load and export every function that
include/tupan_hip.h declares; without a device every entry point fails
loudly (no CPU fallback)."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

from conftest import ROOT, gpu_available

HEADER = os.path.join(ROOT, "include", "tupan_hip.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^[A-Za-z_][\w\s\*]*?\b([A-Za-z_]\w*)\s*\(",
                       text, flags=re.M)
    return sorted(set(n for n in names if n not in ("defined",)))


def exported(path):
    out = subprocess.run(["nm", "-D", "--defined-only", path],
                         stdout=subprocess.PIPE, text=True, check=True).stdout
    return set(line.split()[-1] for line in out.splitlines()
               if " T " in line)


def test_header_lists_reference_and_device_layers():
    names = declared_functions()
    from tupan_amd import _native
    for n in _native.REFERENCE_ABI + _native.DEVICE_ABI:
        assert n in names, n


def test_main_library_exports_every_declared_function():
    from tupan_amd import _native
    syms = exported(_native.LIB_PATH)
    missing = [n for n in declared_functions() if n not in syms]
    assert not missing, missing
    ctypes.CDLL(_native.LIB_PATH)


def test_float_library_exports_reference_names():
    from tupan_amd import _native
    syms = exported(_native.LIB32_PATH)
    for n in _native.REFERENCE_ABI:
        assert n in syms, n
    ctypes.CDLL(_native.LIB32_PATH)


def test_library_built_for_gfx950_only():
    """The embedded device code objects target gfx950 and nothing else."""
    from tupan_amd import _native
    blob = open(_native.LIB_PATH, "rb").read()
    targets = set(re.findall(rb"amdgcn-amd-amdhsa-+(gfx\w+)", blob))
    assert targets == {b"gfx950"}, targets


@pytest.mark.skipif(gpu_available(), reason="checks the no-device path")
def test_no_device_fails_loudly():
    from tupan_amd import _native
    lib = _native.lib()
    cu = ctypes.c_int()
    wg = ctypes.c_int()
    assert lib.tupan_device_info(ctypes.byref(cu), ctypes.byref(wg)) != 0
    assert cu.value == -1
    assert _native.last_error()
    # host-pointer ABI: outputs become NaN, nothing is computed on the CPU
    raw = ctypes.CDLL(_native.LIB_PATH)
    n = 4
    a = [np.ones(n) for _ in range(5)]
    out = np.zeros(n)
    P = ctypes.c_void_p
    raw.phi_kernel.restype = None
    raw.phi_kernel(ctypes.c_ulong(n), *[x.ctypes.data_as(P) for x in a],
                   ctypes.c_ulong(n), *[x.ctypes.data_as(P) for x in a],
                   out.ctypes.data_as(P))
    assert np.all(np.isnan(out))


@pytest.mark.skipif(gpu_available(), reason="checks the no-device path")
def test_extension_without_device_raises():
    from tupan_amd._native import TupanHipError
    from tupan_amd.lib import extensions
    from tupan_amd.particles import ParticleSystem
    ps = ParticleSystem(8, device="cpu")
    ps.mass[...] = 1.0
    ps.rx[...] = torch_arange(8)
    with pytest.raises(TupanHipError):
        extensions.Phi("HIP", "float64").calc(ps, ps)


def torch_arange(n):
    import torch
    return torch.arange(n, dtype=torch.float64)


def test_unknown_backend_and_precision():
    from tupan_amd.lib import extensions
    with pytest.raises(ValueError):
        extensions.Phi("C", "float64")
    with pytest.raises(KeyError):
        extensions.Phi("HIP", "float16")
    assert isinstance(extensions.pnacc, extensions.PNAcc)


def test_new_entry_points_reject_bad_arguments():
    """Argument checks that run before any device work (so they hold with or
    without a GPU): unknown mode, negative near-pair capacity, null outputs
    and null control blocks give TUPAN_ERR_ARG with a message."""
    from tupan_amd import _native
    lib = _native.lib()
    ERR_ARG = -1
    null = None
    arrays = _native.ptr_array([None] * 18)
    assert lib.tupan_sakura_step(0, 99, 4, arrays, null, null, 0.01, 16,
                                 null, null, null) == ERR_ARG
    assert lib.tupan_sakura_step(0, 0, 4, arrays, null, null, 0.01, -1,
                                 null, null, null) == ERR_ARG
    assert "capacity" in _native.last_error()
    fields = _native.ptr_array([None] * 8)
    assert lib.tupan_sakura_graph(4, fields, 4, fields, 0.01, -1, None, 16,
                                  null, null, null) == ERR_ARG
    assert lib.tupan_sakura_graph(4, fields, 4, fields, 0.01, -1,
                                  _native.ptr_array([None] * 6), -5, null,
                                  null, null) == ERR_ARG
    assert lib.tupan_rsq_bias(None) == ERR_ARG
    assert _native.last_error()
