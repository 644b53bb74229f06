"""GPU parity of the pairwise kernels against the reference golden vectors
and the CPU oracle.

Bars (per particle, vector outputs normalised by the particle's vector norm,
scalar outputs relative):
  FP64 (mode fp64):     <= 1e-13
  FP64 (mode ieee64):   bit-exact with the reference C backend when the
                        launch uses one j split (all golden cases), since the
                        operation order, IEEE division and square root are
                        the reference's
  split-FP32:           Acc, AccJerk within 8 FP32 ulps of the sum of
                        contribution magnitudes; Tstep within 8 units of its
                        first-order FP32 error model and 1e-6 relative (the
                        Plummer error budget is in test_gpu_precision.py)
  FP32 (float build):   <= 2e-5 against the reference's float build
PNAcc runs at the reference's own test values (order 7, c = 128) in the
generic tests, and at every order 0..8 and at c = 16 in test_pnacc_orders.
"""
import ctypes

import numpy as np
import pytest

from conftest import vec_rel_err

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

KERNELS = {
    # name: (extension class, scalars, output names, vector groups)
    "phi": ("Phi", (), ("phi",), [(0,)]),
    "acc": ("Acc", (), ("ax", "ay", "az"), [(0, 1, 2)]),
    "acc_jerk": ("AccJerk", (), ("ax", "ay", "az", "jx", "jy", "jz"),
                 [(0, 1, 2), (3, 4, 5)]),
    "snap_crackle": ("SnapCrackle", (), ("sx", "sy", "sz", "cx", "cy", "cz"),
                     [(0, 1, 2), (3, 4, 5)]),
    "tstep": ("Tstep", (1.0 / 64,), ("tstep", "tstepij"), [(0,), (1,)]),
    "nreg_x": ("NREG_X", (1.0 / 64,), ("mrx", "mry", "mrz", "ax", "ay", "az",
                                        "u"), [(0, 1, 2), (3, 4, 5), (6,)]),
    "nreg_v": ("NREG_V", (1.0 / 64,), ("mvx", "mvy", "mvz", "mk"),
               [(0, 1, 2), (3,)]),
    "pnacc": ("PNAcc", (), ("pnax", "pnay", "pnaz"), [(0, 1, 2)]),
}
FIELDS = ("mass", "rx", "ry", "rz", "eps2", "vx", "vy", "vz",
          "ax", "ay", "az", "jx", "jy", "jz")
PREFIXES = (1, 2, 3, 7, 16, 33, 61)  # as written by golden/make_golden.py


def set_clight(order, c):
    from tupan_amd.lib import extensions
    extensions.clight.pn_order = order
    extensions.clight.clight = c


@pytest.fixture(autouse=True)
def pn_parameters():
    """PNAcc reads the module's clight: the reference's test values
    (tupan/tests/test_extensions.py:110-111)."""
    set_clight(7, 128.0)
    yield
    set_clight(0, 128.0)


def device_ps(d, dtype=torch.float64):
    from tupan_amd.particles import ParticleSystem
    from tupan_amd.lib.utils import ctype  # noqa: F401
    ps = ParticleSystem.from_numpy({k: v for k, v in d.items()},
                                   device="cuda:0")
    if dtype == torch.float32:
        for k, v in list(ps.arrays().items()):
            if v.dtype == torch.float64:
                ps.__dict__[k] = v.float()
    return ps


def split_set(g, tag, which, upto=None):
    d = {}
    for k in FIELDS:
        v = g["%s_%s_%s" % (tag, which, k)]
        d[k] = v if upto is None else v[:upto]
    return d


def run_ext(name, prec, mode, ips, jps, scalars):
    from tupan_amd.lib import extensions
    cls = getattr(extensions, KERNELS[name][0])
    ext = cls("HIP", prec, mode=mode)
    res = ext.calc(ips, jps, *scalars)
    torch.cuda.synchronize()
    return [r.detach().cpu().numpy() for r in res]


def assert_close(name, got, g, key, bar):
    outs = KERNELS[name][2]
    ref = [g["%s_%s" % (key, o)] for o in outs]
    for grp in KERNELS[name][3]:
        e = vec_rel_err([got[k] for k in grp], [ref[k] for k in grp])
        assert e <= bar, "%s %s: rel err %.3e > %.1e" % (name, key, e, bar)


@pytest.mark.parametrize("soft", [0, 1])
@pytest.mark.parametrize("name", sorted(KERNELS))
def test_fp64_golden(golden, name, soft):
    g = golden("pairwise_f64.npz")
    tag = "s%d" % soft
    A = device_ps(split_set(g, tag, "A"))
    B = device_ps(split_set(g, tag, "B"))
    sc = KERNELS[name][1]
    for lab, ii, jj in (("AB", A, B), ("BA", B, A), ("AA", A, A)):
        got = run_ext(name, "float64", "fp64", ii, jj, sc)
        assert_close(name, got, g, "%s_%s_%s" % (tag, name, lab), 1e-13)


@pytest.mark.parametrize("soft", [0, 1])
@pytest.mark.parametrize("name", sorted(KERNELS))
def test_ieee64_bitexact(golden, name, soft):
    g = golden("pairwise_f64.npz")
    tag = "s%d" % soft
    A = device_ps(split_set(g, tag, "A"))
    B = device_ps(split_set(g, tag, "B"))
    sc = KERNELS[name][1]
    for lab, ii, jj in (("AB", A, B), ("BA", B, A), ("AA", A, A)):
        got = run_ext(name, "float64", "ieee64", ii, jj, sc)
        for o, r in zip(KERNELS[name][2], got):
            ref = g["%s_%s_%s_%s" % (tag, name, lab, o)]
            assert np.array_equal(r, ref), (name, lab, o,
                                            np.max(np.abs(r - ref)))


@pytest.mark.parametrize("name", sorted(KERNELS))
def test_prefix_sweep(golden, name):
    """A (97) against every prefix B[:k], as the reference's TestCase1
    sweeps j-prefixes (tupan/tests/test_extensions.py:56-83)."""
    g = golden("pairwise_f64.npz")
    sc = KERNELS[name][1]
    for soft in (0, 1):
        tag = "s%d" % soft
        A = device_ps(split_set(g, tag, "A"))
        for k in PREFIXES:
            B = device_ps(split_set(g, tag, "B", k))
            got = run_ext(name, "float64", "fp64", A, B, sc)
            assert_close(name, got, g, "%s_%s_pre%d" % (tag, name, k), 1e-13)


def abs_sums(ips, jps):
    """Per i, the sums over j of |acceleration| and |jerk| contributions
    (the scale an FP32 evaluation's rounding error is proportional to)."""
    dx = ips["rx"][:, None] - jps["rx"][None, :]
    dy = ips["ry"][:, None] - jps["ry"][None, :]
    dz = ips["rz"][:, None] - jps["rz"][None, :]
    ux = ips["vx"][:, None] - jps["vx"][None, :]
    uy = ips["vy"][:, None] - jps["vy"][None, :]
    uz = ips["vz"][:, None] - jps["vz"][None, :]
    r2 = dx * dx + dy * dy + dz * dz
    e2 = ips["eps2"][:, None] + jps["eps2"][None, :]
    with np.errstate(divide="ignore", invalid="ignore"):
        q2 = np.where(r2 > 0, 1.0 / (r2 + e2), 0.0)
    q3 = q2 * np.sqrt(q2) * jps["mass"][None, :]
    r = np.sqrt(r2)
    v = np.sqrt(ux * ux + uy * uy + uz * uz)
    acc = (q3 * r).sum(axis=1)
    jerk = (q3 * (v + 3 * v * r2 * q2)).sum(axis=1)
    return acc, jerk


def test_split32_golden(golden):
    """Split-FP32 error is an FP32 rounding error per contribution: bounded
    by a few FP32 ulps (2^-24 = 6e-8) of the sum of contribution
    magnitudes, whatever cancellation does to the total."""
    g = golden("pairwise_f64.npz")
    for soft in (0, 1):
        tag = "s%d" % soft
        a, b = split_set(g, tag, "A"), split_set(g, tag, "B")
        A, B = device_ps(a), device_ps(b)
        for lab, ii, jj, hi, hj in (("AB", A, B, a, b), ("BA", B, A, b, a),
                                    ("AA", A, A, a, a)):
            sa, sj = abs_sums(hi, hj)
            got = run_ext("acc_jerk", "float64", "split32", ii, jj, ())
            ref = [g["%s_acc_jerk_%s_%s" % (tag, lab, o)]
                   for o in KERNELS["acc_jerk"][2]]
            for grp, scale in (((0, 1, 2), sa), ((3, 4, 5), sj)):
                d = np.sqrt(sum((got[k] - ref[k])**2 for k in grp))
                e = float(np.max(d / scale))
                assert e <= 8 * 2.0**-24, (lab, grp, e)
            got = run_ext("acc", "float64", "split32", ii, jj, ())
            ref = [g["%s_acc_%s_%s" % (tag, lab, o)] for o in ("ax", "ay",
                                                                 "az")]
            d = np.sqrt(sum((got[k] - ref[k])**2 for k in range(3)))
            assert float(np.max(d / sa)) <= 8 * 2.0**-24


def tstep_error_scale(hi, hj, eta, extra_speed=0.0):
    """First-order FP32 error model of split-FP32 Tstep, per i-particle, for
    tstep and tstepij (relative, in units of 2^-24).  Each pair term
    w = w0 - ga*rv carries errors of a few ulps of the magnitudes entering
    it: |w0| with the velocity difference widened by U = |v_i - v_0| +
    |v_j - v_0| (velocities are rounded to FP32 relative to the first
    i-particle v_0), and |ga| |r| (|u| + U); tstep = eta / sqrt(1 + S) then
    halves the relative error of 1 + S.  extra_speed widens U further (a
    bound for a frame velocity that hi[...][0] does not hold)."""
    v0 = np.array([hi[k][0] for k in ("vx", "vy", "vz")])
    vj = np.stack([hj[k] for k in ("vx", "vy", "vz")]).T
    sc_t, sc_m = [], []
    for i in range(len(hi["mass"])):
        d = np.stack([hi[k][i] - hj[k] for k in ("rx", "ry", "rz")])
        u = np.stack([hi[k][i] - hj[k] for k in ("vx", "vy", "vz")])
        vi = np.array([hi[k][i] for k in ("vx", "vy", "vz")])
        U = (np.linalg.norm(vi - v0) + np.linalg.norm(vj - v0, axis=1)
             + extra_speed)
        r2 = (d * d).sum(0)
        mask = r2 > 0
        with np.errstate(all="ignore"):
            q2 = np.where(mask, 1 / (r2 + hi["eps2"][i] + hj["eps2"]), 0)
            phi = (hi["mass"][i] + hj["mass"]) * np.sqrt(q2)
            un = np.sqrt((u * u).sum(0))
            w0 = q2 * (2 * phi + (u * u).sum(0))
            ga = q2 * (2 * q2 * phi + w0) * eta / np.sqrt(w0)
            w = np.where(mask, w0 - ga * (d * u).sum(0), 0)
            mag = np.where(mask, q2 * (2 * phi + (un + U)**2)
                           + np.abs(ga) * np.sqrt(r2) * (un + U), 0)
        sc_t.append(0.5 * mag.sum() / (1 + w.sum()))
        j = int(np.argmax(w))
        sc_m.append(0.5 * mag[j] / (1 + w[j]))
    return np.array(sc_t), np.array(sc_m)


def test_split32_tstep_golden(golden):
    """Split-FP32 Tstep against the reference: within 8 units of the
    first-order FP32 error model (an FP32 emulation of the kernel on these
    sets reaches 1.4 units), and within 1e-6 relative."""
    g = golden("pairwise_f64.npz")
    eta = KERNELS["tstep"][1][0]
    for soft in (0, 1):
        tag = "s%d" % soft
        a, b = split_set(g, tag, "A"), split_set(g, tag, "B")
        A, B = device_ps(a), device_ps(b)
        for lab, ii, jj, hi, hj in (("AB", A, B, a, b), ("BA", B, A, b, a),
                                    ("AA", A, A, a, a)):
            got = run_ext("tstep", "float64", "split32", ii, jj, (eta,))
            scales = tstep_error_scale(hi, hj, eta)
            for k, o in enumerate(("tstep", "tstepij")):
                ref = g["%s_tstep_%s_%s" % (tag, lab, o)]
                e = np.abs(got[k] - ref) / ref
                assert float(e.max()) <= 1e-6, (lab, o, float(e.max()))
                units = float(np.max(e / (scales[k] * 2.0**-24)))
                assert units <= 8, (lab, o, units)


@pytest.mark.parametrize("name", sorted(KERNELS))
def test_fp32_golden(golden, name):
    g = golden("pairwise_f32.npz")
    sc = KERNELS[name][1]
    for soft in (0, 1):
        tag = "s%d" % soft
        A = device_ps(split_set(g, tag, "A"), torch.float32)
        B = device_ps(split_set(g, tag, "B"), torch.float32)
        for lab, ii, jj in (("AB", A, B), ("BA", B, A), ("AA", A, A)):
            got = run_ext(name, "float32", "fp64", ii, jj, sc)
            assert_close(name, got, g, "%s_%s_%s" % (tag, name, lab), 2e-5)


@pytest.mark.parametrize("soft", [0, 1])
def test_pnacc_orders(golden, soft):
    """Every PN order 0..8 at c = 128 and order 7 at c = 16: FP64 within
    1e-13, ieee64 bit-exact, the float build within 2e-5."""
    g = golden("pairwise_f64.npz")
    g32 = golden("pairwise_f32.npz")
    tag = "s%d" % soft
    A = device_ps(split_set(g, tag, "A"))
    B = device_ps(split_set(g, tag, "B"))
    A32 = device_ps(split_set(g32, tag, "A"), torch.float32)
    B32 = device_ps(split_set(g32, tag, "B"), torch.float32)
    cases = [("o%d" % o, o, 128.0) for o in range(9)] + [("c16", 7, 16.0)]
    for key, order, c in cases:
        set_clight(order, c)
        for lab, ii, jj, i32, j32 in (("AB", A, B, A32, B32),
                                      ("AA", A, A, A32, A32)):
            full = "%s_pnacc_%s_%s" % (tag, key, lab)
            ref = [g["%s_%s" % (full, o)] for o in ("pnax", "pnay", "pnaz")]
            got = run_ext("pnacc", "float64", "ieee64", ii, jj, ())
            for r, e in zip(got, ref):
                assert np.array_equal(r, e), (key, lab)
            got = run_ext("pnacc", "float64", "fp64", ii, jj, ())
            if order < 2:
                assert all(not np.any(r) for r in got), (key, lab)
                continue
            assert_close("pnacc", got, g, full, 1e-13)
            got = run_ext("pnacc", "float32", "fp64", i32, j32, ())
            assert_close("pnacc", got, g32, full, 2e-5)


def test_split32_pnacc(golden):
    """Split-FP32 PNAcc against the reference at orders 2, 4, 5, 6, 7
    (c = 128) and 7 (c = 16): FP32 pair terms on FP64 inputs, within 2e-6
    of the PN acceleration per particle (its size relative to the Newtonian
    acceleration is printed; on these sets v/c is 0.04 to 0.08)."""
    g = golden("pairwise_f64.npz")
    errs = {}
    for soft in (0, 1):
        tag = "s%d" % soft
        A = device_ps(split_set(g, tag, "A"))
        B = device_ps(split_set(g, tag, "B"))
        cases = [("o%d" % o, o, 128.0) for o in (2, 4, 5, 6, 7)]
        cases.append(("c16", 7, 16.0))
        for key, order, c in cases:
            set_clight(order, c)
            for lab, ii, jj in (("AB", A, B), ("AA", A, A)):
                full = "%s_pnacc_%s_%s" % (tag, key, lab)
                ref = [g["%s_%s" % (full, o)] for o in ("pnax", "pnay",
                                                        "pnaz")]
                got = run_ext("pnacc", "float64", "split32", ii, jj, ())
                an = [g["%s_acc_%s_%s" % (tag, lab, o)]
                      for o in ("ax", "ay", "az")]
                d = np.sqrt(sum((x - y)**2 for x, y in zip(got, ref)))
                en = float(np.max(d / np.sqrt(sum(x * x for x in an))))
                errs[(tag, key, lab)] = (vec_rel_err(got, ref), en)
    for k, (e, en) in sorted(errs.items()):
        print("split32 pnacc %s: %.3e of PN acc, %.3e of Newtonian" % (
            "/".join(k), e, en))
    worst = max(errs.items(), key=lambda kv: kv[1][0])
    assert worst[1][0] <= 2e-6, worst


def test_host_pointer_abi(golden):
    """The reference-named host-pointer entry points (libtupan.h ABI) on
    numpy arrays, called through ctypes as cffi would call the reference."""
    from tupan_amd import _native
    lib = ctypes.CDLL(_native.LIB_PATH)
    g = golden("pairwise_f64.npz")
    A = split_set(g, "s1", "A")
    B = split_set(g, "s1", "B")
    P = ctypes.c_void_p

    def p(a):
        return np.ascontiguousarray(a).ctypes.data_as(P)
    ni, nj = len(A["mass"]), len(B["mass"])
    out = [np.zeros(ni) for _ in range(6)]
    lib.acc_jerk_kernel.restype = None
    lib.acc_jerk_kernel(ctypes.c_ulong(ni), *[p(A[k]) for k in FIELDS[:8]],
                        ctypes.c_ulong(nj), *[p(B[k]) for k in FIELDS[:8]],
                        *[p(o) for o in out])
    assert_close("acc_jerk", out, g, "s1_acc_jerk_AB", 1e-13)
    out2 = [np.zeros(ni) for _ in range(2)]
    lib.tstep_kernel.restype = None
    lib.tstep_kernel(ctypes.c_ulong(ni), *[p(A[k]) for k in FIELDS[:8]],
                     ctypes.c_ulong(nj), *[p(B[k]) for k in FIELDS[:8]],
                     ctypes.c_double(1.0 / 64), *[p(o) for o in out2])
    assert_close("tstep", out2, g, "s1_tstep_AB", 1e-13)
    lib.pnacc_kernel.restype = None
    inv = [ctypes.c_double(128.0**-k) for k in range(1, 8)]
    out3 = [np.zeros(ni) for _ in range(3)]
    lib.pnacc_kernel(ctypes.c_ulong(ni), *[p(A[k]) for k in FIELDS[:8]],
                     ctypes.c_ulong(nj), *[p(B[k]) for k in FIELDS[:8]],
                     ctypes.c_ulong(7), *inv, *[p(o) for o in out3])
    assert_close("pnacc", out3, g, "s1_pnacc_AB", 1e-13)
    # the other four names of libtupan.h: phi, snap_crackle, nreg_X, nreg_V
    out4 = [np.zeros(ni)]
    lib.phi_kernel.restype = None
    lib.phi_kernel(ctypes.c_ulong(ni), *[p(A[k]) for k in FIELDS[:5]],
                   ctypes.c_ulong(nj), *[p(B[k]) for k in FIELDS[:5]],
                   *[p(o) for o in out4])
    assert_close("phi", out4, g, "s1_phi_AB", 1e-13)
    out5 = [np.zeros(ni) for _ in range(6)]
    lib.snap_crackle_kernel.restype = None
    lib.snap_crackle_kernel(ctypes.c_ulong(ni), *[p(A[k]) for k in FIELDS],
                            ctypes.c_ulong(nj), *[p(B[k]) for k in FIELDS],
                            *[p(o) for o in out5])
    assert_close("snap_crackle", out5, g, "s1_snap_crackle_AB", 1e-13)
    dt = KERNELS["nreg_x"][1][0]
    out6 = [np.zeros(ni) for _ in range(7)]
    lib.nreg_Xkernel.restype = None
    lib.nreg_Xkernel(ctypes.c_ulong(ni), *[p(A[k]) for k in FIELDS[:8]],
                     ctypes.c_ulong(nj), *[p(B[k]) for k in FIELDS[:8]],
                     ctypes.c_double(dt), *[p(o) for o in out6])
    assert_close("nreg_x", out6, g, "s1_nreg_x_AB", 1e-13)
    vfields = ("mass", "vx", "vy", "vz", "ax", "ay", "az")
    out7 = [np.zeros(ni) for _ in range(4)]
    lib.nreg_Vkernel.restype = None
    lib.nreg_Vkernel(ctypes.c_ulong(ni), *[p(A[k]) for k in vfields],
                     ctypes.c_ulong(nj), *[p(B[k]) for k in vfields],
                     ctypes.c_double(KERNELS["nreg_v"][1][0]),
                     *[p(o) for o in out7])
    assert_close("nreg_v", out7, g, "s1_nreg_v_AB", 1e-13)


def test_host_pointer_abi_f32(golden):
    from tupan_amd import _native
    lib = ctypes.CDLL(_native.LIB32_PATH)
    g = golden("pairwise_f32.npz")
    A = split_set(g, "s0", "A")
    B = split_set(g, "s0", "B")
    P = ctypes.c_void_p

    def p(a):
        return np.ascontiguousarray(a, dtype=np.float32).ctypes.data_as(P)
    ni, nj = len(A["mass"]), len(B["mass"])
    out = [np.zeros(ni, np.float32) for _ in range(3)]
    lib.acc_kernel.restype = None
    lib.acc_kernel(ctypes.c_uint(ni), *[p(A[k]) for k in FIELDS[:5]],
                   ctypes.c_uint(nj), *[p(B[k]) for k in FIELDS[:5]],
                   *[p(o) for o in out])
    assert_close("acc", out, g, "s0_acc_AB", 2e-5)
    lib.pnacc_kernel.restype = None
    inv = [ctypes.c_float(128.0**-k) for k in range(1, 8)]
    out3 = [np.zeros(ni, np.float32) for _ in range(3)]
    lib.pnacc_kernel(ctypes.c_uint(ni), *[p(A[k]) for k in FIELDS[:8]],
                     ctypes.c_uint(nj), *[p(B[k]) for k in FIELDS[:8]],
                     ctypes.c_uint(7), *inv, *[p(o) for o in out3])
    assert_close("pnacc", out3, g, "s0_pnacc_AB", 2e-5)


def test_numpy_staging_through_extension(golden):
    """Host numpy arrays through the HIP backend (staged copies, outputs
    copied back by map_buffers), as the reference's OpenCL backend does."""
    from tupan_amd.lib import extensions

    class HostPS(object):
        def __init__(self, d):
            self.__dict__.update({k: np.array(v) for k, v in d.items()})
            self.n = len(d["mass"])

        def register_auxiliary_attribute(self, name, sctype):
            self.__dict__[name] = np.zeros(self.n)
    g = golden("pairwise_f64.npz")
    A = HostPS(split_set(g, "s1", "A"))
    B = HostPS(split_set(g, "s1", "B"))
    res = extensions.AccJerk("HIP", "float64").calc(A, B)
    assert all(isinstance(r, np.ndarray) for r in res)
    assert res[0] is A.ax
    assert_close("acc_jerk", list(res), g, "s1_acc_jerk_AB", 1e-13)


def test_empty_and_degenerate():
    from tupan_amd.lib import extensions
    rng = np.random.default_rng(3)
    d = {k: rng.random(5) for k in FIELDS}
    d["eps2"] = np.zeros(5)
    ps = device_ps(d)
    empty = device_ps({k: v[:0] for k, v in d.items()})
    # nj == 0: sums are zero, Tstep gives eta
    r = extensions.AccJerk("HIP", "float64").calc(ps, empty)
    for x in r:
        assert torch.all(x == 0)
    r = extensions.Tstep("HIP", "float64").calc(ps, empty, 0.25)
    assert torch.all(r[0] == 0.25) and torch.all(r[1] == 0.25)
    # ni == 0: nothing to do
    r = extensions.Acc("HIP", "float64").calc(empty, ps)
    assert all(x.numel() == 0 for x in r)
    # a lone particle with itself: masked, exactly zero
    one = device_ps({k: v[:1] for k, v in d.items()})
    r = extensions.AccJerk("HIP", "float64").calc(one, one)
    for x in r:
        assert float(x[0]) == 0.0
    # PNAcc: nj == 0, ni == 0, a lone particle, orders 0 and 1 -> zeros
    for mode in ("fp64", "ieee64", "split32"):
        pn = extensions.PNAcc("HIP", "float64", mode=mode)
        for order in (0, 1, 7):
            set_clight(order, 16.0)
            for x in pn.calc(ps, empty):
                assert torch.all(x == 0)
            assert all(x.numel() == 0 for x in pn.calc(empty, ps))
            for x in pn.calc(one, one):
                assert float(x[0]) == 0.0
            if order < 2:
                for x in pn.calc(ps, ps):
                    assert torch.all(x == 0)
    r = extensions.Phi("HIP", "float64").calc(one, one)
    assert float(r[0][0]) == 0.0


@pytest.mark.parametrize("mode", ["fp64", "split32"])
def test_large_random_rows_vs_oracle(mode):
    """N = 65536 uniform cube: a sample of rows against the oracle, plus
    momentum conservation over all rows (sum m_i a_i = 0 up to rounding)."""
    from oracle import Oracle
    from tupan_amd.lib import extensions
    rng = np.random.default_rng(11)
    n = 65536
    d = {k: rng.random(n) for k in ("rx", "ry", "rz", "vx", "vy", "vz")}
    d["mass"] = np.full(n, 1.0 / n)
    d["eps2"] = np.full(n, 0.5e-8)
    ps = device_ps(d)
    res = extensions.AccJerk("HIP", "float64", mode=mode).calc(ps, ps)
    got = [r.detach().cpu().numpy() for r in res]
    rows = rng.choice(n, 64, replace=False)
    sub = {k: v[rows] for k, v in d.items()}
    ref = Oracle("float64").acc_jerk(sub, d)
    bar = 1e-13 if mode == "fp64" else 1e-6
    e = vec_rel_err([got[k][rows] for k in range(3)], ref[:3])
    assert e <= bar, e
    fx = float(np.sum(d["mass"] * got[0]))
    amag = float(np.sum(d["mass"] * np.abs(got[0])))
    assert abs(fx) <= (1e-12 if mode == "fp64" else 1e-6) * amag


@pytest.mark.parametrize("mode", ["fp64", "split32"])
def test_block_step_size_salpeter_rows_vs_oracle(mode):
    """The block-step workload's size: Plummer N = 262144 with a Salpeter
    mass function (slope -2.35, masses 0.1-10 normalised to total mass 1),
    eps = 1e-4.  Acc of the whole set, and of an uneven contiguous i-subset
    (one block-step level, a view into the same arrays) against all j:
    sampled rows of both against the oracle (FP64 within 1e-13 of |a|;
    split-FP32 within 8 FP32 ulps of the row's sum of contribution
    magnitudes) and momentum conservation over all rows."""
    from oracle import Oracle
    from tupan_amd.ics.imf import IMF
    from tupan_amd.ics.plummer import make_plummer
    from tupan_amd.lib import extensions
    n = 262144
    ps = make_plummer(n, 1e-4, IMF.power_law(-2.35, 0.1, 10.0), seed=5,
                      device="cuda:0")
    d = ps.to_numpy()
    acc = extensions.Acc("HIP", "float64", mode=mode)
    full = [r.detach().cpu().numpy().copy() for r in acc.calc(ps, ps)]
    lo, hi = 100003, 100003 + 37717
    part = [r.detach().cpu().numpy().copy()
            for r in acc.calc(ps[lo:hi], ps)]
    keys = ("mass", "rx", "ry", "rz", "eps2", "vx", "vy", "vz")
    dj = {k: d[k] for k in keys}
    rng = np.random.default_rng(5)
    for got, rows, off in ((full, rng.choice(n, 24, replace=False), 0),
                           (part, lo + rng.choice(hi - lo, 12,
                                                  replace=False), lo)):
        di = {k: d[k][rows] for k in keys}
        ref = Oracle("float64").acc(di, dj)
        g = [got[k][rows - off] for k in range(3)]
        if mode == "fp64":
            e = vec_rel_err(g, ref[:3])
            print("N = %d %s rows at %d: %.3e of |a|" % (n, mode, off, e))
            assert e <= 1e-13, (off, e)
        else:
            sa, _ = abs_sums(di, dj)
            err = np.sqrt(sum((g[k] - ref[k])**2 for k in range(3)))
            e = float(np.max(err / sa))
            print("N = %d %s rows at %d: %.3e of sum |contrib|"
                  % (n, mode, off, e))
            assert e <= 8 * 2.0**-24, (off, e)
    m = d["mass"]
    for k in range(3):
        f = float(np.sum(m * full[k]))
        amag = float(np.sum(m * np.abs(full[k])))
        assert abs(f) <= (1e-12 if mode == "fp64" else 1e-6) * amag, k


@pytest.mark.parametrize("mode", ["fp64", "split32"])
def test_block_step_workload_tstep_rows_vs_oracle(mode):
    """Tstep (eta = 0.02) on the block-step workload itself: Plummer
    N = 262144, Salpeter masses (slope -2.35 on [0.1, 10], total mass 1),
    eps = 1e-4, seed 1.  The whole set and an uneven contiguous i-subset
    (one level, a view into the same arrays) against all j; sampled rows
    against the oracle: FP64 within 1e-13 relative; split-FP32 within 1e-6
    relative and within 8 units of the first-order FP32 error model
    (tstep_error_scale, with the frame velocity bounded by the largest
    speed, since the workgroup's first i-particle is not known here)."""
    from oracle import Oracle
    from tupan_amd.ics.imf import IMF
    from tupan_amd.ics.plummer import make_plummer
    from tupan_amd.lib import extensions
    n, eta = 262144, 0.02
    ps = make_plummer(n, 1e-4, IMF.power_law(-2.35, 0.1, 10.0), seed=1,
                      device="cuda:0")
    d = ps.to_numpy()
    ext = extensions.Tstep("HIP", "float64", mode=mode)
    full = [r.detach().cpu().numpy().copy() for r in ext.calc(ps, ps, eta)]
    lo, hi = 70001, 70001 + 23459
    part = [r.detach().cpu().numpy().copy()
            for r in ext.calc(ps[lo:hi], ps, eta)]
    keys = ("mass", "rx", "ry", "rz", "eps2", "vx", "vy", "vz")
    dj = {k: d[k] for k in keys}
    vmax = float(np.max(np.sqrt(d["vx"]**2 + d["vy"]**2 + d["vz"]**2)))
    rng = np.random.default_rng(7)
    for got, rows, off in ((full, rng.choice(n, 24, replace=False), 0),
                           (part, lo + rng.choice(hi - lo, 12,
                                                  replace=False), lo)):
        di = {k: d[k][rows] for k in keys}
        ref = Oracle("float64").tstep(di, dj, eta)
        for k, o in enumerate(("tstep", "tstepij")):
            e = np.abs(got[k][rows - off] - ref[k]) / ref[k]
            print("N = %d %s %s rows at %d: max rel %.3e"
                  % (n, mode, o, off, e.max()))
            if mode == "fp64":
                assert float(e.max()) <= 1e-13, (o, off)
            else:
                assert float(e.max()) <= 1e-6, (o, off)
        if mode == "split32":
            # frame velocity v0 unknown: |v_i - v0| + |v_j - v0| is at most
            # |v_i| + |v_j| + 2 vmax; the model takes v0 = 0 and the
            # widened speeds through its U term
            hi_ = dict(di)
            for c in ("vx", "vy", "vz"):
                hi_[c] = np.concatenate([[0.0], di[c]])
            for c in ("mass", "rx", "ry", "rz", "eps2"):
                hi_[c] = np.concatenate([[di[c][0]], di[c]])
            sc = tstep_error_scale(hi_, dj, eta, extra_speed=2 * vmax)
            for k in range(2):
                e = np.abs(got[k][rows - off] - ref[k]) / ref[k]
                units = float(np.max(e / (sc[k][1:] * 2.0**-24)))
                print("  error model units (%s): %.2f" % (k, units))
                assert units <= 8, (k, off, units)


def test_split_combine_order():
    """The j-split partial sums are combined in a fixed order.  With the
    reference-order FP64 pair kernel each split's partial sum equals the
    oracle's sum over that j-range bit for bit, so the GPU result can be
    rebuilt on the host: for N = 1024 cut into 4 splits of 256 (shape
    64x1x64x4; ieee64 never cuts below 256) the combine of small sets (8
    lanes per i, lane totals added pairwise) gives (p0 + p1) + (p2 + p3),
    which differs from the left-to-right sum in the last bit for some i."""
    import ctypes as C
    from oracle import Oracle
    from tupan_amd import _native
    from tupan_amd.lib import extensions
    from tupan_amd.tune import KERNEL_IDS, SLOTS
    lib = _native.lib()
    kid, slot = KERNEL_IDS["acc_jerk"], SLOTS["ieee64"]
    old_shape, old_wg = C.c_int(), C.c_int()
    assert lib.tupan_tune_get(kid, slot, 0, C.byref(old_shape), C.byref(old_wg)) == 0
    rng = np.random.default_rng(5)
    n = 1024
    d = {k: rng.random(n) for k in ("rx", "ry", "rz", "vx", "vy", "vz")}
    d["mass"] = rng.uniform(0.5, 1.5, n) / n
    d["eps2"] = np.full(n, 1e-6)
    ps = device_ps(d)
    try:
        assert lib.tupan_tune_set(kid, slot, 0, 5, 16) == 0
        res = extensions.AccJerk("HIP", "float64", mode="ieee64").calc(ps, ps)
        got = [r.detach().cpu().numpy() for r in res]
    finally:
        lib.tupan_tune_set(kid, slot, 0, old_shape.value, old_wg.value)
    orc = Oracle("float64")
    parts = [orc.acc_jerk(d, {k: v[s:s + 256] for k, v in d.items()})
             for s in range(0, n, 256)]
    for k in range(6):
        p = [q[k] for q in parts]
        tree = (p[0] + p[1]) + (p[2] + p[3])
        seq = ((p[0] + p[1]) + p[2]) + p[3]
        assert np.array_equal(got[k], tree), k
        print("component %d: tree and left-to-right differ for %d of %d"
              % (k, int(np.sum(tree != seq)), n))


@pytest.mark.parametrize("name", ["phi", "acc", "acc_jerk", "tstep"])
def test_ieee64_uniform_softening_bitexact(name):
    """Every particle with the same nonzero eps2: the tile engine forms the
    pair softening once per tile (e2 + e2) instead of per pair; the result
    is the reference's bit for bit (one j-split, reference order), with
    i and j sets of different sizes and a duplicated particle."""
    from oracle import Oracle
    rng = np.random.default_rng(17)
    d = {}
    for n, tag in ((203, "i"), (241, "j")):
        x = {k: rng.normal(size=n) for k in ("rx", "ry", "rz", "vx", "vy", "vz")}
        x["mass"] = rng.uniform(0.5, 2.0, n) / n
        x["eps2"] = np.full(n, 3.7e-5)
        d[tag] = x
    for k in d["i"]:
        d["j"][k][5] = d["i"][k][9]          # one particle in both sets
    ii, jj = device_ps(d["i"]), device_ps(d["j"])
    sc = KERNELS[name][1]
    got = run_ext(name, "float64", "ieee64", ii, jj, sc)
    ref = getattr(Oracle("float64"), name)(d["i"], d["j"], *sc)
    for o, r, e in zip(KERNELS[name][2], got, ref):
        assert np.array_equal(r, e), (name, o, np.max(np.abs(r - e)))


def test_full_prefix_sweep_ieee64_vs_oracle():
    """The reference's TestCase1 sweep in full (tupan/tests/test_extensions.py
    :28-83): 256 particles of two member types (the set is two concatenated
    halves), masses U[0, 1), eps2 = 0, positions and velocities U[0, 10)
    (seeded here), i = all 256 against every j-prefix 1..256 and the swapped
    roles, for the eight pairwise kernels; the reference-order FP64 mode
    must equal the oracle bit for bit at every size."""
    from oracle import Oracle
    rng = np.random.default_rng(29)
    n = 256
    halves = []
    for m in (n - n // 2, n // 2):
        h = {"mass": rng.random(m), "eps2": np.zeros(m)}
        for k in ("rx", "ry", "rz", "vx", "vy", "vz", "ax", "ay", "az",
                  "jx", "jy", "jz"):
            h[k] = 10.0 * rng.random(m)
        halves.append(h)
    d = {k: np.concatenate([h[k] for h in halves]) for k in halves[0]}
    orc = Oracle("float64")
    inv = (1.0 / 128.0,)
    pn = (7,) + tuple(inv[0] ** k for k in range(1, 8))
    for name in sorted(KERNELS):
        sc = KERNELS[name][1]
        osc = pn if name == "pnacc" else sc
        for jdx in range(1, n + 1):
            sub = {k: v[:jdx] for k, v in d.items()}
            for hi, hj in ((d, sub), (sub, d)):
                # fresh device sets: outputs land in the i-set's arrays
                # (Acc writes ax, which NREG_V and SnapCrackle read)
                ii = device_ps(hi)
                jj = ii if hj is hi else device_ps(hj)
                got = run_ext(name, "float64", "ieee64", ii, jj, sc)
                ref = getattr(orc, name)(hi, hj, *osc)
                for o, r, e in zip(KERNELS[name][2], got, ref):
                    assert np.array_equal(r, e, equal_nan=True), (name, jdx, o)


@pytest.mark.parametrize("name", sorted(KERNELS))
def test_small_set_kernel_vs_oracle(name):
    """Sets of a few hundred to two thousand j-particles take the one-launch
    small-set kernel (64 lanes per i, fixed shuffle-tree combine): every
    kernel against the oracle on 700 particles of two member types (eps2 of
    each type 0 and 1e-4), i = j and a 300-particle i-set against the 700,
    FP64 within 1e-13 of the FP64 oracle (relative to the vector's norm),
    the float build within 1e-4 of the oracle's float build (its sums run in
    another order), split32 (where it exists) within 1e-5 of the FP64
    oracle; the precision of each mode is pinned elsewhere, this checks the
    small-set path of every kernel."""
    from oracle import Oracle
    rng = np.random.default_rng(31)
    n = 700
    d = {k: rng.normal(size=n) for k in ("rx", "ry", "rz", "vx", "vy", "vz",
                                         "ax", "ay", "az", "jx", "jy", "jz")}
    d["mass"] = rng.uniform(0.5, 2.0, n) / n
    d["eps2"] = np.where(np.arange(n) < n // 2, 0.0, 1e-4)
    sub = {k: v[:300] for k, v in d.items()}
    sc = KERNELS[name][1]
    osc = ((7,) + tuple((1.0 / 128.0) ** k for k in range(1, 8))
           if name == "pnacc" else sc)
    for hi, hj in ((d, d), (sub, d)):
        modes = [("float64", "fp64", 1e-13), ("float32", "fp64", 1e-4)]
        if name in ("acc", "acc_jerk", "tstep", "pnacc"):
            modes.append(("float64", "split32", 1e-5))
        for prec, mode, bar in modes:
            ref = getattr(Oracle(prec), name)(hi, hj, *osc)
            dt = torch.float32 if prec == "float32" else torch.float64
            ii = device_ps(hi, dt)
            jj = ii if hj is hi else device_ps(hj, dt)
            got = run_ext(name, prec, mode, ii, jj, sc)
            for grp in KERNELS[name][3]:
                e = vec_rel_err([np.asarray(got[k], np.float64) for k in grp],
                                [ref[k] for k in grp])
                assert e <= bar, (name, prec, mode, len(hi["mass"]), grp, e)


@pytest.mark.parametrize("n", [600, 3000])
def test_split32_fp32_overflow_falls_back_to_fp64(n):
    """Unsoftened close pairs overflow the split-FP32 pair terms (m / R^3 at
    R = 1e-15 is ~1e42, beyond FP32): the kernels evaluate every i whose
    sums come out non-finite again in FP64, so split32 Acc, AccJerk, Tstep
    and PNAcc give the FP64 path's values there instead of inf or NaN.  One
    pair approaches and one recedes (the Tstep pair term's sign case).
    N = 600 runs the one-launch small-set kernel (the lanes of an i each
    redo their own j), N = 3000 the tile engine (an i redoes its j-split).
    Close-pair rows: within 1e-12 of the FP64 path (both FP64, summed in a
    different order); every other row finite and within 1e-5."""
    from tupan_amd.lib import extensions
    rng = np.random.default_rng(21)
    d = {k: rng.random(n) for k in ("rx", "ry", "rz", "vx", "vy", "vz")}
    d["mass"] = np.full(n, 1.0 / n) * (1 + rng.random(n))
    d["eps2"] = np.zeros(n)
    # rows 0..3 lie near row 0, the frame of the first workgroup: 0-1 at
    # 1e-15 approaching, 2-3 at 1e-15 receding, 1e-6 from 0-1
    x0 = (d["rx"][0], d["ry"][0], d["rz"][0])
    d["rx"][1] = x0[0] + 1e-15
    d["ry"][1], d["rz"][1] = x0[1], x0[2]
    d["rx"][2], d["ry"][2], d["rz"][2] = x0[0], x0[1] + 1e-6, x0[2]
    d["rx"][3], d["ry"][3], d["rz"][3] = x0[0] + 1e-15, x0[1] + 1e-6, x0[2]
    d["vx"][0], d["vx"][1] = 1.0, -1.0
    d["vx"][2], d["vx"][3] = -1.0, 1.0
    ps = device_ps(d)
    close = [0, 1, 2, 3]
    set_clight(7, 128.0)
    for name, cls, sc in (("acc", "Acc", ()), ("acc_jerk", "AccJerk", ()),
                          ("tstep", "Tstep", (1.0 / 64,)),
                          ("pnacc", "PNAcc", ())):
        ref = [r.detach().cpu().numpy() for r in getattr(extensions, cls)(
            "HIP", "float64", mode="fp64").calc(ps, ps, *sc)]
        got = [r.detach().cpu().numpy() for r in getattr(extensions, cls)(
            "HIP", "float64", mode="split32").calc(ps, ps, *sc)]
        for grp in KERNELS[name][3]:
            g = [got[k] for k in grp]
            r = [ref[k] for k in grp]
            assert np.all(np.isfinite(np.stack(g))), name
            e_close = vec_rel_err([x[close] for x in g], [x[close] for x in r])
            rest = np.ones(n, bool)
            rest[close] = False
            e_rest = vec_rel_err([x[rest] for x in g], [x[rest] for x in r])
            print("%s n=%d: close rows %.2e, others %.2e" % (name, n, e_close,
                                                             e_rest))
            assert e_close <= 1e-12, (name, e_close)
            assert e_rest <= 1e-5, (name, e_rest)
