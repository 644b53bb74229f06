"""Unmasked uniform-softening tiles of the Acc and AccJerk, split-FP32 and
FP64 (ops.h: self_safe, unmask_allowed; pairwise.hip: the coincidence
scan), on Plummer N = 16384 (eps = 1e-4, 2^28 pairs: the AccJerk call runs
the scan).

  * On a Plummer sphere the unmasked sums equal the masked ones bit for bit:
    tupan_pairwise (scan, unmasked tiles) against tupan_pairwise_noalloc
    (no scan, masked tiles, the same launch plan).
  * Two particles 1e-9 apart (r^2 > 0, but below half an FP32 ulp of the
    pair softening, so R^2 = e2 in FP32) with different velocities: the
    reference keeps the pair (smoothing.h: r^2 > 0), and so do the
    unmasked tiles; the pair's term dominates both particles' jerks.
    Their rows match the oracle's within the split error.
  * The same two particles at one position: the reference drops the pair
    (r^2 = 0); the scan finds it and the call keeps the mask, so no jerk
    term appears.  Their rows match the oracle's.  Both scan forms run:
    the system against itself (insertion alone) and a copy of it as i-set
    (insertion of j, then a probe per i).
"""
import numpy as np
import pytest

from conftest import vec_rel_err

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

N = 16384


def plummer():
    from oracle_backend import install_hip
    from tupan_amd.ics.plummer import make_plummer
    install_hip("float64")
    return make_plummer(N, 1e-4, ("equalmass",), seed=2, device="cuda:0")


def place_pair(ps, dx):
    """Particle 1 at particle 0's position + (dx, 0, 0), with a velocity
    0.5 away from particle 0's."""
    for k in ("rx", "ry", "rz"):
        getattr(ps, k)[1] = getattr(ps, k)[0]
    ps.rx[1] = ps.rx[0] + dx
    ps.vx[1] = ps.vx[0] + 0.3
    ps.vy[1] = ps.vy[0] - 0.4
    ps.vz[1] = ps.vz[0]
    torch.cuda.synchronize()


# tolerances against the oracle's rows: split-FP32 error, FP64 rounding
TOL = {"split32": (1e-5, 1e-4), "fp64": (1e-12, 1e-11)}


def split_rows(ps, kernel, rows, iset="same", mode="split32"):
    """Rows of the call with the system as i- and j-set ("same": one set,
    the scan's insertion alone), or a copy as i-set ("copy": the scan
    probes the j table with every i)."""
    from tupan_amd.lib import extensions
    k = getattr(extensions, kernel)("HIP", "float64", mode=mode)
    ips = ps if iset == "same" else ps.copy()
    res = k.calc(ips, ps)
    return [r.detach().cpu().numpy()[rows] for r in res]


def oracle_rows(ps, kernel, rows):
    from oracle import Oracle
    host = ps.to_numpy()
    o = Oracle("float64")
    fn = o.acc_jerk if kernel == "AccJerk" else o.acc
    return fn({k: v[rows] for k, v in host.items()}, host)


@pytest.mark.parametrize("mode", ["split32", "fp64"])
def test_unmasked_equals_masked_on_plummer(mode):
    from tupan_amd import _native
    ps = plummer()
    fields = [ps.mass, ps.rx, ps.ry, ps.rz, ps.eps2, ps.vx, ps.vy, ps.vz]
    ip = _native.ptr_array([f.data_ptr() for f in fields])
    lib = _native.lib()
    outs = {}
    for name in ("tupan_pairwise", "tupan_pairwise_noalloc"):
        print("%s AccJerk N = %d through %s" % (mode, N, name), flush=True)
        out = [torch.full((N,), np.nan, dtype=torch.float64, device="cuda:0")
               for _ in range(6)]
        op = _native.ptr_array([o.data_ptr() for o in out])
        rc = getattr(lib, name)(2, _native.MODES[mode], N, ip, N, ip,
                                op, 0.0, None)
        _native.check(rc, name)
        torch.cuda.synchronize()
        outs[name] = [o.cpu().numpy() for o in out]
    for a, b in zip(outs["tupan_pairwise"], outs["tupan_pairwise_noalloc"]):
        assert np.all(np.isfinite(a))
        assert np.array_equal(a, b)


@pytest.mark.parametrize("kernel,iset,mode", [("AccJerk", "same", "split32"),
                                              ("AccJerk", "copy", "split32"),
                                              ("Acc", "same", "split32"),
                                              ("AccJerk", "same", "fp64"),
                                              ("Acc", "same", "fp64")])
def test_tiny_separation_pair_is_kept(kernel, iset, mode):
    ps = plummer()
    place_pair(ps, 1e-9)
    print("%s %s, pair 1e-9 apart, N = %d, i-set %s" % (mode, kernel, N, iset),
          flush=True)
    rows = np.array([0, 1, 2])
    got = split_rows(ps, kernel, rows, iset, mode)
    ref = oracle_rows(ps, kernel, rows)
    tol = TOL[mode][0]
    assert vec_rel_err(got[:3], ref[:3]) <= tol
    if kernel == "AccJerk":
        assert vec_rel_err(got[3:], ref[3:]) <= tol
        # the pair's m |dv| / eps^3 term (~1.5e7) dominates the jerks
        assert np.hypot(ref[3][0], ref[4][0]) > 1e6


@pytest.mark.parametrize("iset,mode", [("same", "split32"),
                                       ("copy", "split32"),
                                       ("same", "fp64"),
                                       ("copy", "fp64")])
def test_coincident_pair_keeps_the_mask(iset, mode):
    ps = plummer()
    place_pair(ps, 0.0)
    print("%s AccJerk, coincident pair, N = %d, i-set %s" % (mode, N, iset),
          flush=True)
    rows = np.array([0, 1, 2])
    got = split_rows(ps, "AccJerk", rows, iset, mode)
    ref = oracle_rows(ps, "AccJerk", rows)
    assert np.all(np.abs(np.stack(ref[3:])) < 1e4)
    assert vec_rel_err(got[:3], ref[:3]) <= TOL[mode][0]
    assert vec_rel_err(got[3:], ref[3:]) <= TOL[mode][1]
