"""ctypes binding of libtupan_hip.so (include/tupan_hip.h).

The library is built in-tree by ``python -m tupan_amd.csrc.build`` into
``tupan_amd/_lib``.  There is no fallback: if the library is missing or no
HIP device is present, calls raise.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_lib", "libtupan_hip.so")
LIB32_PATH = os.path.join(HERE, "_lib", "libtupan_hip_f32.so")

# Status codes and modes (include/tupan_hip.h).
OK = 0
MODE_FP64 = 0
MODE_SPLIT32 = 1
MODE_IEEE64 = 2
MODES = {"fp64": MODE_FP64, "split32": MODE_SPLIT32, "ieee64": MODE_IEEE64}

(K_PHI, K_ACC, K_ACC_JERK, K_SNAP_CRACKLE, K_TSTEP, K_NREG_X, K_NREG_V,
 K_PNACC) = range(8)

_lib = None

_P = ctypes.c_void_p
_S = ctypes.c_size_t
_PP = ctypes.POINTER(ctypes.c_void_p)


class TupanHipError(RuntimeError):
    pass


class SolveLog(ctypes.Structure):
    """tupan_solve_log (include/tupan_hip.h)."""
    _fields_ = [("features", ctypes.c_void_p), ("target", ctypes.c_void_p),
                ("iterations", ctypes.c_void_p), ("capacity", ctypes.c_int)]


def _declare(lib):
    lib.tupan_last_error.restype = ctypes.c_char_p
    lib.tupan_last_error.argtypes = []
    lib.tupan_version.restype = ctypes.c_int
    lib.tupan_device_info.restype = ctypes.c_int
    lib.tupan_device_info.argtypes = [ctypes.POINTER(ctypes.c_int)] * 2
    lib.tupan_timing.restype = ctypes.c_int
    lib.tupan_timing.argtypes = [ctypes.c_int]
    lib.tupan_timing_read.restype = ctypes.c_int
    lib.tupan_timing_read.argtypes = [ctypes.POINTER(ctypes.c_double),
                                      ctypes.POINTER(ctypes.c_int)]
    lib.tupan_release_workspaces.restype = ctypes.c_int
    lib.tupan_release_workspaces.argtypes = []
    lib.tupan_timing_read_each.restype = ctypes.c_int
    lib.tupan_timing_read_each.argtypes = [ctypes.POINTER(ctypes.c_double),
                                           ctypes.c_int,
                                           ctypes.POINTER(ctypes.c_int)]
    for fn in ("tupan_tune_set",):
        getattr(lib, fn).restype = ctypes.c_int
        getattr(lib, fn).argtypes = [ctypes.c_int] * 5
    IP = ctypes.POINTER(ctypes.c_int)
    lib.tupan_tune_get.restype = ctypes.c_int
    lib.tupan_tune_get.argtypes = [ctypes.c_int] * 3 + [IP, IP]
    lib.tupan_rsq_bias.restype = ctypes.c_int
    lib.tupan_rsq_bias.argtypes = [ctypes.POINTER(ctypes.c_double)]
    lib.tupan_tune_shape.restype = ctypes.c_int
    lib.tupan_tune_shape.argtypes = [ctypes.c_int] * 2 + [IP] * 4
    lib.tupan_pairwise.restype = ctypes.c_int
    lib.tupan_pairwise.argtypes = [ctypes.c_int, ctypes.c_int, _S, _PP, _S, _PP,
                                   _PP, ctypes.c_double, _P]
    lib.tupan_pairwise_noalloc.restype = ctypes.c_int
    lib.tupan_pairwise_noalloc.argtypes = lib.tupan_pairwise.argtypes
    lib.tupan_pairwise_f32.restype = ctypes.c_int
    lib.tupan_pairwise_f32.argtypes = [ctypes.c_int, ctypes.c_int, _S, _PP, _S,
                                       _PP, _PP, ctypes.c_float, _P]
    lib.tupan_hermite_step.restype = ctypes.c_int
    lib.tupan_hermite_step.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                       _S, _PP, _P, _P, ctypes.c_double, _P]
    lib.tupan_sia_step.restype = ctypes.c_int
    lib.tupan_sia_step.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                   _S, _PP, _P, _P, ctypes.c_double,
                                   ctypes.c_int, _P, _P, _P, _P]
    lib.tupan_sia_levels_create.restype = ctypes.c_int
    lib.tupan_sia_levels_create.argtypes = [
        ctypes.c_int, _PP, ctypes.c_int, _P, ctypes.c_int, ctypes.c_int,
        ctypes.c_double, ctypes.c_int, _P, _P, ctypes.c_int, _P, _P,
        ctypes.POINTER(ctypes.c_void_p)]
    lib.tupan_sia_levels_destroy.restype = ctypes.c_int
    lib.tupan_sia_levels_destroy.argtypes = [_P]
    lib.tupan_sia_levels_set.restype = ctypes.c_int
    lib.tupan_sia_levels_set.argtypes = [_P, ctypes.c_double, ctypes.c_double,
                                         ctypes.c_double, ctypes.c_int]
    lib.tupan_sia_levels_enqueue.restype = ctypes.c_int
    lib.tupan_sia_levels_enqueue.argtypes = [_P, ctypes.c_int, ctypes.c_int,
                                             _P]
    lib.tupan_sia_levels_state.restype = ctypes.c_int
    lib.tupan_sia_levels_state.argtypes = [_P, _P, _P, _P]
    lib.tupan_pnacc.restype = ctypes.c_int
    lib.tupan_pnacc.argtypes = [ctypes.c_int, _S, _PP, _S, _PP, ctypes.c_ulong,
                                ctypes.POINTER(ctypes.c_double), _PP, _P]
    lib.tupan_pnacc_f32.restype = ctypes.c_int
    lib.tupan_pnacc_f32.argtypes = [ctypes.c_int, _S, _PP, _S, _PP,
                                    ctypes.c_ulong,
                                    ctypes.POINTER(ctypes.c_float), _PP, _P]
    lib.tupan_sakura.restype = ctypes.c_int
    lib.tupan_sakura.argtypes = ([_S] + [_P] * 8 + [_S] + [_P] * 8
                                 + [ctypes.c_double, ctypes.c_long]
                                 + [_P] * 6 + [_P, _P, _P])
    lib.tupan_ensemble_create.restype = ctypes.c_int
    lib.tupan_ensemble_create.argtypes = [ctypes.c_int, _P, ctypes.c_int,
                                          ctypes.POINTER(ctypes.c_void_p)]
    lib.tupan_ensemble_destroy.restype = ctypes.c_int
    lib.tupan_ensemble_destroy.argtypes = [_P]
    lib.tupan_ensemble_info.restype = ctypes.c_int
    lib.tupan_ensemble_info.argtypes = [_P] + [ctypes.POINTER(ctypes.c_int)] * 3
    lib.tupan_ensemble_set_guess.restype = ctypes.c_int
    lib.tupan_ensemble_set_guess.argtypes = [_P, _P, _P]
    lib.tupan_ensemble_sakura.restype = ctypes.c_int
    lib.tupan_ensemble_sakura.argtypes = ([_P] + [_P] * 8
                                          + [ctypes.c_double, ctypes.c_long]
                                          + [_P] * 6 + [_P, _P, _P])
    lib.tupan_ensemble_phi.restype = ctypes.c_int
    lib.tupan_ensemble_phi.argtypes = [_P] + [_P] * 5 + [_P, _P]
    lib.tupan_kepler_pairs.restype = ctypes.c_int
    lib.tupan_kepler_pairs.argtypes = ([_S, _P] + [_P] * 8
                                       + [ctypes.c_double, _P, _P])
    lib.tupan_sakura_ex.restype = ctypes.c_int
    lib.tupan_sakura_ex.argtypes = [_S, _PP, _S, _PP, ctypes.c_double,
                                    ctypes.c_long, _PP, ctypes.c_int, _P,
                                    ctypes.POINTER(SolveLog), _P,
                                    ctypes.POINTER(ctypes.c_int), _P, _P]
    lib.tupan_sakura_step.restype = ctypes.c_int
    lib.tupan_sakura_step.argtypes = [ctypes.c_int, ctypes.c_int, _S, _PP, _P,
                                      _P, ctypes.c_double, ctypes.c_int, _P,
                                      _P, _P]
    lib.tupan_sakura_graph.restype = ctypes.c_int
    lib.tupan_sakura_graph.argtypes = [_S, _PP, _S, _PP, ctypes.c_double,
                                       ctypes.c_long, _PP, ctypes.c_int, _P,
                                       _P, _P]
    lib.tupan_sia_subtree.restype = ctypes.c_int
    lib.tupan_sia_subtree.argtypes = ([ctypes.c_int, _PP, ctypes.c_int, _P,
                                       ctypes.c_int, ctypes.c_int,
                                       ctypes.c_double, ctypes.c_double,
                                       ctypes.c_int, _P, _P, ctypes.c_int, _P,
                                       _P, _P, _P, _P, _P])
    lib.tupan_mlp_sizes.restype = ctypes.c_int
    lib.tupan_mlp_sizes.argtypes = [ctypes.POINTER(ctypes.c_int)] * 3
    lib.tupan_mlp_grad.restype = ctypes.c_int
    lib.tupan_mlp_grad.argtypes = [ctypes.c_int] + [_P] * 7
    lib.tupan_mlp_forward.restype = ctypes.c_int
    lib.tupan_mlp_forward.argtypes = [ctypes.c_int] + [_P] * 5
    lib.tupan_mlp_adam.restype = ctypes.c_int
    lib.tupan_mlp_adam.argtypes = ([_P] * 4 + [ctypes.c_float] * 4
                                   + [ctypes.c_int, _P])
    lib.tupan_mlp_pack.restype = ctypes.c_int
    lib.tupan_mlp_pack.argtypes = [_P] * 4
    lib.tupan_mlp_infer.restype = ctypes.c_int
    lib.tupan_mlp_infer.argtypes = [ctypes.c_int] + [_P] * 4
    return lib


def lib():
    """The loaded library (loads on first use; raises if it is not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise TupanHipError(
                "libtupan_hip.so is not built at %s; run "
                "`python -m tupan_amd.csrc.build`" % LIB_PATH)
        _lib = _declare(ctypes.CDLL(LIB_PATH))
        from .tune import apply
        apply(_lib)
    return _lib


def last_error():
    return lib().tupan_last_error().decode(errors="replace")


def check(rc, what):
    if rc != OK:
        raise TupanHipError("%s failed (%d): %s" % (what, rc, last_error()))


def ptr_array(ptrs):
    arr = (ctypes.c_void_p * max(1, len(ptrs)))()
    for k, p in enumerate(ptrs):
        arr[k] = p
    return arr


def timing(enable):
    check(lib().tupan_timing(1 if enable else 0), "tupan_timing")


def timing_read():
    """(summed kernel milliseconds, launches) since timing(True)."""
    ms = ctypes.c_double()
    cnt = ctypes.c_int()
    check(lib().tupan_timing_read(ctypes.byref(ms), ctypes.byref(cnt)),
          "tupan_timing_read")
    return ms.value, cnt.value


def timing_read_each():
    """Milliseconds of each recorded launch interval since timing(True)."""
    cnt = ctypes.c_int()
    check(lib().tupan_timing_read_each(None, 0, ctypes.byref(cnt)),
          "tupan_timing_read_each")
    out = (ctypes.c_double * max(1, cnt.value))()
    check(lib().tupan_timing_read_each(out, cnt.value, ctypes.byref(cnt)),
          "tupan_timing_read_each")
    return list(out)[:cnt.value]


def release_workspaces():
    """Free the library's scratch workspaces on the current device (after a
    device synchronisation); HIP graphs captured earlier that contain
    library launches must be captured again."""
    check(lib().tupan_release_workspaces(), "tupan_release_workspaces")


def device_info():
    cu = ctypes.c_int()
    wg = ctypes.c_int()
    rc = lib().tupan_device_info(ctypes.byref(cu), ctypes.byref(wg))
    check(rc, "tupan_device_info")
    return cu.value, wg.value


def rsq_bias():
    """tupan_rsq_bias: (raw mean error on [1, 2), on [2, 4), refined on
    [1, 2), on [2, 4), compiled-in bias) of v_rsq_f32 on this device."""
    out = (ctypes.c_double * 5)()
    check(lib().tupan_rsq_bias(out), "tupan_rsq_bias")
    return tuple(out)


# Names exported by the shared libraries (used by tests that check the ABI).
REFERENCE_ABI = ("phi_kernel", "acc_kernel", "acc_jerk_kernel",
                 "snap_crackle_kernel", "tstep_kernel", "sakura_kernel",
                 "kepler_solver_kernel", "nreg_Xkernel", "nreg_Vkernel",
                 "pnacc_kernel")
DEVICE_ABI = ("tupan_last_error", "tupan_version", "tupan_device_info",
              "tupan_tune_set", "tupan_tune_get", "tupan_tune_shape",
              "tupan_rsq_bias",
              "tupan_timing", "tupan_timing_read", "tupan_timing_read_each",
              "tupan_phi", "tupan_acc", "tupan_acc_jerk", "tupan_snap_crackle",
              "tupan_tstep", "tupan_pairwise", "tupan_pairwise_f32",
              "tupan_pairwise_noalloc", "tupan_release_workspaces",
              "tupan_sakura", "tupan_kepler_pairs", "tupan_host_pairwise_f32",
              "tupan_host_sakura_f32", "tupan_host_kepler_f32",
              "tupan_ensemble_create", "tupan_ensemble_destroy",
              "tupan_ensemble_info", "tupan_ensemble_set_guess",
              "tupan_ensemble_sakura",
              "tupan_ensemble_phi", "tupan_sakura_ex", "tupan_mlp_sizes",
              "tupan_mlp_grad", "tupan_mlp_forward", "tupan_mlp_adam",
              "tupan_mlp_pack", "tupan_mlp_infer", "tupan_sia_subtree",
              "tupan_pnacc", "tupan_pnacc_f32", "tupan_host_pnacc_f32",
              "tupan_hermite_step", "tupan_sakura_graph",
              "tupan_sakura_step", "tupan_sia_step",
              "tupan_sia_levels_create", "tupan_sia_levels_destroy",
              "tupan_sia_levels_set", "tupan_sia_levels_enqueue",
              "tupan_sia_levels_state")
