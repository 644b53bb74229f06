"""Force precision at the full 65536-particle Plummer size (eps = 1e-4,
seed 1) through size-independent properties:

  * FP64 fast path vs the reference-order IEEE path (bit-exact per pair with
    the reference C backend): <= 1e-13 per particle, normalised by |a_i|;
  * FP64 rows against the CPU oracle on a sample of particles;
  * split-FP32 against FP64: the per-particle relative force error
    e_i = |da_i|/|a_i|.  The 1e-7 force-error target is judged on its root
    mean square over the particles (the usual N-body force-accuracy
    measure), which must be <= 1e-7; beside it the median <= 5e-8, the
    99.9th percentile <= 5e-7 and the maximum <= 5e-6 (measured on an
    MI355X: median 1.4e-8, 99.9th percentile 2.7e-7, maximum 1.7e-6; the
    largest values sit where the net force nearly cancels); the jerk's
    median <= 1e-6 (the jerk is a derivative and cancels harder);
  * split-FP32 forces carry no common scale error (virial within 2e-9 of
    FP64's), the condition for |dE/E| <= 1e-10 over one N-body time;
  * split-FP32 Tstep against FP64: relative error median <= 1e-7, 99.9th
    percentile <= 1e-6, block-step levels unchanged for all but <= 1e-4 of
    the particles;
  * Newton's third law: sum m_i a_i = 0 to rounding in both paths.
"""
import numpy as np
import pytest

from conftest import vec_rel_err

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def plummer64k():
    from oracle_backend import install_hip
    from tupan_amd.ics.plummer import make_plummer
    install_hip("float64")
    return make_plummer(65536, 1e-4, ("equalmass",), seed=1, device="cuda:0")


def accjerk(ps, mode):
    from tupan_amd.lib import extensions
    res = extensions.AccJerk("HIP", "float64", mode=mode).calc(ps, ps)
    return [r.detach().cpu().numpy().copy() for r in res]


def per_particle(got, ref, grp):
    g = np.stack([got[k] for k in grp])
    r = np.stack([ref[k] for k in grp])
    return np.sqrt(((g - r)**2).sum(0)) / np.sqrt((r * r).sum(0))


def test_fp64_vs_reference_order(plummer64k):
    fast = accjerk(plummer64k, "fp64")
    ieee = accjerk(plummer64k, "ieee64")
    assert vec_rel_err(fast[:3], ieee[:3]) <= 1e-13
    assert vec_rel_err(fast[3:], ieee[3:]) <= 1e-12


def test_fp64_rows_vs_oracle(plummer64k):
    from oracle import Oracle
    host = plummer64k.to_numpy()
    got = accjerk(plummer64k, "fp64")
    rows = np.random.default_rng(1).choice(65536, 48, replace=False)
    ref = Oracle("float64").acc_jerk({k: v[rows] for k, v in host.items()},
                                     host)
    assert vec_rel_err([g[rows] for g in got[:3]], ref[:3]) <= 1e-13
    assert vec_rel_err([g[rows] for g in got[3:]], ref[3:]) <= 1e-12


def test_split32_force_error(plummer64k):
    fast = accjerk(plummer64k, "fp64")
    split = accjerk(plummer64k, "split32")
    ea = per_particle(split, fast, (0, 1, 2))
    ej = per_particle(split, fast, (3, 4, 5))
    rms = float(np.sqrt(np.mean(ea * ea)))
    print("split32 |da|/|a|: rms %.3e median %.3e p99.9 %.3e max %.3e"
          % (rms, np.median(ea), np.percentile(ea, 99.9), ea.max()))
    print("split32 |dj|/|j|: median %.3e p99.9 %.3e max %.3e"
          % (np.median(ej), np.percentile(ej, 99.9), ej.max()))
    assert rms <= 1e-7
    assert np.median(ea) <= 5e-8
    assert np.percentile(ea, 99.9) <= 5e-7
    assert ea.max() <= 5e-6
    assert np.median(ej) <= 1e-6


def test_split32_force_law_unbiased(plummer64k):
    """No common scale error in the split forces: the virial
    W = sum_i m_i r_i . a_i (the potential energy, for 1/r pair forces) of
    split32 equals FP64's to 2e-9, for Acc and AccJerk.  A common relative
    force error beta conserves K + (1 + beta) U instead of K + U, so |dE/E|
    follows beta * dU / E: on a 16384-particle Plummer sphere U changes by
    1.6 % over one N-body time, and beta = 1e-8 (the size of the v_rsq_f32
    bias or of rounding r^2 before adding e2) shows as |dE/E| ~ 5e-10.
    Measured after the fix: |beta| ~ 1e-10."""
    from tupan_amd.lib import extensions
    h = {k: v.cpu().numpy() for k, v in plummer64k.arrays().items()}
    m, rx, ry, rz = h["mass"], h["rx"], h["ry"], h["rz"]

    def virial(a):
        return float(np.sum(m * (rx * a[0] + ry * a[1] + rz * a[2])))

    for name in ("Acc", "AccJerk"):
        w = {}
        for mode in ("fp64", "split32"):
            res = getattr(extensions, name)("HIP", "float64", mode=mode).calc(
                plummer64k, plummer64k)
            w[mode] = virial([r.detach().cpu().numpy() for r in res[:3]])
        beta = w["split32"] / w["fp64"] - 1
        print("%s split32 virial scale error %.3e" % (name, beta))
        assert abs(beta) <= 2e-9


def test_rsq_bias_calibration():
    """The bias the split kernels compensate is the device's: v_rsq_f32's
    mean relative error over every float in [1, 4), averaged over the two
    exponent parities, equals the compiled-in constant to 1e-11 (the
    hardware estimate is deterministic), and one Newton step leaves
    |mean| <= 1e-10."""
    from tupan_amd import _native
    raw0, raw1, ref0, ref1, b = _native.rsq_bias()
    print("v_rsq_f32 mean error: %.4e %.4e, refined %.2e %.2e, compiled %.4e"
          % (raw0, raw1, ref0, ref1, b))
    assert abs(0.5 * (raw0 + raw1) - b) <= 1e-11
    assert max(abs(ref0), abs(ref1)) <= 1e-10


def test_split32_tstep_error_and_levels(plummer64k):
    """Split-FP32 Tstep against FP64 on the Plummer state: relative error
    median <= 1e-7, 99.9th percentile <= 1e-6, and the power-of-two level a
    block-step scheme assigns (floor(log2(tstep / eta))) differs for at most
    one particle in 10^4."""
    from tupan_amd.lib import extensions
    eta = 0.02
    out = {}
    for mode in ("fp64", "split32"):
        res = extensions.Tstep("HIP", "float64", mode=mode).calc(
            plummer64k, plummer64k, eta)
        out[mode] = [r.detach().cpu().numpy().copy() for r in res]
    for k, name in enumerate(("tstep", "tstepij")):
        ref, got = out["fp64"][k], out["split32"][k]
        e = np.abs(got - ref) / ref
        print("split32 %s: median %.3e p99.9 %.3e max %.3e"
              % (name, np.median(e), np.percentile(e, 99.9), e.max()))
        assert np.median(e) <= 1e-7
        assert np.percentile(e, 99.9) <= 1e-6
        lv_ref = np.floor(np.log2(ref / eta))
        lv_got = np.floor(np.log2(got / eta))
        flips = int(np.sum(lv_ref != lv_got))
        print("split32 %s: %d level changes of %d" % (name, flips, ref.size))
        assert flips <= ref.size // 10000


def test_momentum_conservation(plummer64k):
    m = plummer64k.mass.cpu().numpy()
    for mode, bar in (("fp64", 1e-13), ("split32", 1e-7)):
        res = accjerk(plummer64k, mode)
        for comp in res[:3]:
            tot = abs(np.sum(m * comp))
            scale = np.sum(np.abs(m * comp))
            assert tot <= bar * scale, (mode, tot / scale)
