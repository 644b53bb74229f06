"""Pinning of the CPU oracle (oracle/oracle.c).

Each oracle function is checked against
  * the golden vectors written by the reference C backend compiled from its
    own sources (tests/golden/make_golden.py): bit-exact, FP64 and FP32;
  * closed forms and invariants that a plausible slip (a dropped term, a
    sign, a swapped operand) would break: two-body acceleration, jerk,
    potential and time step; Newton's third law (sum m_i a_i = 0,
    sum m_i j_i = 0); masked self pairs; the Kepler solver's exact period,
    energy and angular-momentum conservation, the Lagrange identity
    f g' - g f' = 1, time reversal; Sakura's far-pair leapfrog branch.
"""
import numpy as np
import pytest

from oracle import Oracle, SPECS

FIELDS = ("mass", "rx", "ry", "rz", "eps2", "vx", "vy", "vz",
          "ax", "ay", "az", "jx", "jy", "jz")
PREFIXES = (1, 2, 3, 7, 16, 33, 61)


def sets(g, tag, n=None):
    a = {k: g["%s_A_%s" % (tag, k)] for k in FIELDS}
    b = {k: g["%s_B_%s" % (tag, k)] for k in FIELDS}
    if n is not None:
        b = {k: v[:n] for k, v in b.items()}
    return a, b


@pytest.mark.parametrize("prec", ["float64", "float32"])
@pytest.mark.parametrize("name", ["phi", "acc", "acc_jerk", "snap_crackle",
                                  "tstep", "nreg_x", "nreg_v"])
def test_pairwise_bitexact_vs_reference(golden, prec, name):
    g = golden("pairwise_%s.npz" % ("f64" if prec == "float64" else "f32"))
    orc = Oracle(prec)
    sc = (1.0 / 64,) if name in ("tstep", "nreg_x", "nreg_v") else ()
    for soft in (0, 1):
        tag = "s%d" % soft
        a, b = sets(g, tag)
        for lab, ii, jj in (("AB", a, b), ("BA", b, a), ("AA", a, a)):
            res = orc.pairwise(name, ii, jj, *sc)
            for o, r in zip(SPECS[name][3], res):
                ref = g["%s_%s_%s_%s" % (tag, name, lab, o)]
                assert np.array_equal(r, ref), (name, lab, o)
        for k in PREFIXES:
            _, bk = sets(g, tag, k)
            res = orc.pairwise(name, a, bk, *sc)
            for o, r in zip(SPECS[name][3], res):
                assert np.array_equal(r, g["%s_%s_pre%d_%s" % (tag, name, k,
                                                               o)])


def test_sakura_bitexact_vs_reference(golden):
    g = golden("sakura_f64.npz")
    orc = Oracle("float64")
    for lab in ("r0", "r1", "nb"):
        d = {k: g["%s_%s" % (lab, k)] for k in FIELDS[:8]}
        for flag in (-2, -1, 0, 1, 2):
            for dt in (1.0 / 64, 0.5):
                res = orc.sakura(d, d, dt, flag)
                for o, r in zip(SPECS["sakura"][3], res):
                    ref = g["%s_f%d_dt%g_%s" % (lab, flag, dt, o)]
                    assert np.array_equal(r, ref), (lab, flag, dt, o)


def test_kepler_bitexact_vs_reference(golden):
    g = golden("kepler_f64.npz")
    orc = Oracle("float64")
    for c in range(int(g["ncases"])):
        d = {k: g["c%d_%s" % (c, k)] for k in FIELDS[:8]}
        res = orc.kepler(d, float(g["c%d_dt" % c]))
        for o, r in zip(("rx", "ry", "rz", "vx", "vy", "vz"), res):
            assert np.array_equal(r, g["c%d_out_%s" % (c, o)]), (c, o)


def two_body(m0=0.3, m1=0.7, e2=0.0):
    d = {"mass": np.array([m0, m1]), "eps2": np.array([e2, e2]),
         "rx": np.array([0.1, 1.3]), "ry": np.array([-0.2, 0.4]),
         "rz": np.array([0.05, -0.6]), "vx": np.array([0.3, -0.1]),
         "vy": np.array([0.0, 0.5]), "vz": np.array([-0.2, 0.1])}
    return d


@pytest.mark.parametrize("e2", [0.0, 0.01])
def test_two_body_closed_forms(e2):
    """Acc, AccJerk, Phi, Tstep of a pair against the closed forms of
    acc_jerk_kernel_common.h:7-56 and tstep_kernel_common.h:7-60 written
    out by hand: a_0 = -m_1 r/R^3, j_0 = -m_1 (v/R^3 - 3 (r.v) r/R^5),
    phi_0 = -m_1/R with R^2 = r^2 + e2_0 + e2_1 and r = r_0 - r_1."""
    d = two_body(e2=e2)
    orc = Oracle("float64")
    r = np.array([d["rx"][0] - d["rx"][1], d["ry"][0] - d["ry"][1],
                  d["rz"][0] - d["rz"][1]])
    v = np.array([d["vx"][0] - d["vx"][1], d["vy"][0] - d["vy"][1],
                  d["vz"][0] - d["vz"][1]])
    R2 = r @ r + 2 * e2
    R = np.sqrt(R2)
    m1 = d["mass"][1]
    a0 = -m1 * r / R**3
    j0 = -m1 * (v / R**3 - 3 * (r @ v) * r / R**5)
    ax, ay, az, jx, jy, jz = orc.acc_jerk(d, d)
    np.testing.assert_allclose([ax[0], ay[0], az[0]], a0, rtol=1e-14)
    np.testing.assert_allclose([jx[0], jy[0], jz[0]], j0, rtol=1e-13)
    (phi,) = orc.phi(d, d)
    np.testing.assert_allclose(phi[0], -m1 / R, rtol=1e-15)
    # Newton's third law
    m = d["mass"]
    np.testing.assert_allclose(m[0] * ax[0] + m[1] * ax[1], 0, atol=1e-15)
    np.testing.assert_allclose(m[0] * jy[0] + m[1] * jy[1], 0, atol=1e-15)
    # Tstep: w2 = (v^2 + 2M/R)/R^2, gamma = (w2 + 2M/R^3)/R^2 * eta/sqrt(w2),
    # w2 -= gamma (r.v); dt = eta/sqrt(1 + w2) for both outputs (one pair)
    eta = 0.03
    M = m.sum()
    w2 = (v @ v + 2 * M / R) / R2
    gam = (w2 + 2 * M / R**3) / R2 * eta / np.sqrt(w2)
    w2 = w2 - gam * (r @ v)
    dt_a, dt_b = orc.tstep(d, d, eta)
    np.testing.assert_allclose(dt_a[0], eta / np.sqrt(1 + w2), rtol=1e-14)
    np.testing.assert_allclose(dt_b[0], eta / np.sqrt(1 + max(w2, 0)),
                               rtol=1e-14)


def test_invariants_and_self_mask(golden):
    g = golden("pairwise_f64.npz")
    orc = Oracle("float64")
    a, _ = sets(g, "s1")
    m = a["mass"]
    ax, ay, az, jx, jy, jz = orc.acc_jerk(a, a)
    for comp in (ax, ay, az, jx, jy, jz):
        scale = np.sum(np.abs(m * comp))
        assert abs(np.sum(m * comp)) <= 1e-13 * scale
    # a lone particle sees nothing (mask on r2 > 0), even with softening
    one = {k: v[:1] for k, v in a.items()}
    for name in ("phi", "acc", "acc_jerk", "snap_crackle"):
        for x in orc.pairwise(name, one, one):
            assert x[0] == 0.0
    dt_a, dt_b = orc.tstep(one, one, 0.5)
    assert dt_a[0] == 0.5 and dt_b[0] == 0.5


def test_snap_crackle_matches_finite_difference_of_jerk():
    """Snap is the time derivative of jerk along the motion: with the
    positions, velocities and accelerations advanced consistently, a
    centred difference of AccJerk's jerk reproduces SnapCrackle's snap."""
    rng = np.random.default_rng(4)
    n = 6
    d = {k: rng.normal(0, 1, n) for k in ("rx", "ry", "rz", "vx", "vy",
                                         "vz")}
    d["mass"] = rng.random(n) + 0.5
    d["eps2"] = np.full(n, 0.05)
    orc = Oracle("float64")
    ax, ay, az, jx, jy, jz = orc.acc_jerk(d, d)
    d.update(ax=ax, ay=ay, az=az, jx=jx, jy=jy, jz=jz)
    sx, sy, sz, cx, cy, cz = orc.snap_crackle(d, d)
    h = 1e-4

    def state(t):
        e = dict(d)
        for c, a, j, s in (("x", ax, jx, sx), ("y", ay, jy, sy),
                           ("z", az, jz, sz)):
            e["r" + c] = d["r" + c] + d["v" + c] * t + a * t**2 / 2 \
                + j * t**3 / 6 + s * t**4 / 24
            e["v" + c] = d["v" + c] + a * t + j * t**2 / 2 + s * t**3 / 6
        return e
    jp = orc.acc_jerk(state(h), state(h))
    jm = orc.acc_jerk(state(-h), state(-h))
    snap_fd = [(jp[k] - jm[k]) / (2 * h) for k in (3, 4, 5)]
    for got, fd in zip((sx, sy, sz), snap_fd):
        np.testing.assert_allclose(got, fd, rtol=2e-5, atol=1e-6)


def rel_state(m=1.0):
    p = np.array([1.0, 0.0, 0.0])
    v = np.array([0.0, np.sqrt(m) * 0.8, 0.1])
    return p, v


def test_kepler_period_energy_angular_momentum():
    orc = Oracle("float64")
    m = 1.0
    p0, v0 = rel_state(m)
    r0 = np.linalg.norm(p0)
    alpha = v0 @ v0 - 2 * m / r0
    a = m / abs(alpha)
    T = 2 * np.pi * a * np.sqrt(a / m)
    for dt in (0.1, 0.37 * T, T, 3 * T + 0.2):
        p1, v1 = orc.kepler_rel(dt, m, 0.0, p0, v0)
        e0 = 0.5 * v0 @ v0 - m / r0
        e1 = 0.5 * v1 @ v1 - m / np.linalg.norm(p1)
        assert abs(e1 - e0) <= 1e-12 * abs(e0)
        np.testing.assert_allclose(np.cross(p1, v1), np.cross(p0, v0),
                                   rtol=1e-12, atol=1e-14)
    p1, v1 = orc.kepler_rel(T, m, 0.0, p0, v0)
    np.testing.assert_allclose(p1, p0, atol=1e-11)
    np.testing.assert_allclose(v1, v0, atol=1e-11)


@pytest.mark.parametrize("vy", [0.5, 1.0, 1.3, 2.5])
def test_kepler_time_reversal_and_lagrange_identity(vy):
    orc = Oracle("float64")
    m = 1.0
    p0 = np.array([1.0, 0.2, 0.0])
    v0 = np.array([0.1, vy, -0.05])
    dt = 0.7
    p1, v1 = orc.kepler_rel(dt, m, 0.0, p0, v0)
    p2, v2 = orc.kepler_rel(-dt, m, 0.0, p1, v1)
    np.testing.assert_allclose(p2, p0, rtol=1e-11, atol=1e-12)
    np.testing.assert_allclose(v2, v0, rtol=1e-11, atol=1e-12)
    # with p1 = f p0 + g v0, v1 = fd p0 + gd v0: f gd - g fd = 1
    A = np.stack([p0, v0], axis=1)
    f, g = np.linalg.lstsq(A, p1, rcond=None)[0]
    fd, gd = np.linalg.lstsq(A, v1, rcond=None)[0]
    assert abs(f * gd - g * fd - 1) <= 1e-10


def test_sakura_far_pair_is_one_leapfrog_step():
    """r > 64 m / v^2 takes the leapfrog branch (sakura_kernel_common.h:
    111-114): drift dt/2, kick, drift dt/2."""
    orc = Oracle("float64")
    d = {"mass": np.array([1e-3, 1e-3]), "eps2": np.zeros(2),
         "rx": np.array([0.0, 10.0]), "ry": np.zeros(2), "rz": np.zeros(2),
         "vx": np.array([0.0, 0.0]), "vy": np.array([1.0, 0.0]),
         "vz": np.zeros(2)}
    dt = 0.25
    res = orc.sakura(d, d, dt, 1)
    # pair (0, 1): r = r0 - r1, v = v0 - v1, M = 2e-3; flag +1: K(dt) then
    # r -= v dt; contribution to particle 0 is (m1/M) (state1 - state0)
    r = np.array([-10.0, 0.0, 0.0])
    v = np.array([0.0, 1.0, 0.0])
    M = 2e-3
    rh = r + v * dt / 2
    a = -M * rh / np.linalg.norm(rh)**3
    v1 = v + a * dt
    r1 = rh + v1 * dt / 2
    r1 = r1 - v1 * dt
    mu = 0.5
    np.testing.assert_allclose([res[0][0], res[1][0], res[2][0]],
                               mu * (r1 - r), rtol=1e-13, atol=1e-16)
    np.testing.assert_allclose([res[3][0], res[4][0], res[5][0]],
                               mu * (v1 - v), rtol=1e-13, atol=1e-18)


def test_sakura_flag_zero_is_identity(golden):
    g = golden("sakura_f64.npz")
    for dt in (1.0 / 64, 0.5):
        for o in SPECS["sakura"][3]:
            assert np.all(g["nb_f0_dt%g_%s" % (dt, o)] == 0)


def test_counters_count_solves_and_leapfrogs():
    orc = Oracle("float64")
    rng = np.random.default_rng(8)
    n = 20
    d = {k: rng.normal(0, 1, n) for k in ("rx", "ry", "rz", "vx", "vy",
                                         "vz")}
    d["mass"] = np.full(n, 1.0 / n)
    d["eps2"] = np.zeros(n)
    orc.counters_reset()
    orc.sakura(d, d, 0.01, -1)
    solves, iters, retries, leap, fails = orc.counters()
    assert solves + leap == n * (n - 1)
    assert iters >= solves and fails == 0


def test_nreg_two_body_closed_form():
    """NREG X and V for one pair against their closed forms: the separation
    drifted by dt, the softened 1/R and 1/R^3, m_i m_j / R and
    m_i m_j |v + a dt|^2; self pairs contribute m_j * 0 and nothing else."""
    orc = Oracle("float64")
    d = {"mass": np.array([0.3, 0.7]), "eps2": np.array([0.01, 0.02]),
         "rx": np.array([0.5, -0.2]), "ry": np.array([0.1, 0.4]),
         "rz": np.array([-0.3, 0.2]), "vx": np.array([0.2, -0.1]),
         "vy": np.array([0.0, 0.3]), "vz": np.array([0.1, 0.1]),
         "ax": np.array([0.05, -0.02]), "ay": np.array([0.0, 0.01]),
         "az": np.array([-0.03, 0.04])}
    dt = 0.125
    mrx, mry, mrz, ax, ay, az, u = orc.nreg_x(d, d, dt)
    r = np.array([d["rx"][0] - d["rx"][1], d["ry"][0] - d["ry"][1],
                  d["rz"][0] - d["rz"][1]])
    v = np.array([d["vx"][0] - d["vx"][1], d["vy"][0] - d["vy"][1],
                  d["vz"][0] - d["vz"][1]])
    rd = r + v * dt
    R = np.sqrt(rd @ rd + 0.03)
    m0, m1 = d["mass"]
    np.testing.assert_allclose([mrx[0], mry[0], mrz[0]], m1 * rd, rtol=1e-14)
    np.testing.assert_allclose([ax[0], ay[0], az[0]], -m1 * rd / R**3,
                               rtol=1e-13)
    np.testing.assert_allclose(u[0], m0 * m1 / R, rtol=1e-14)
    np.testing.assert_allclose([mrx[1], mry[1], mrz[1]], -m0 * rd,
                               rtol=1e-14)
    mvx, mvy, mvz, k = orc.nreg_v(d, d, dt)
    a = np.array([d["ax"][0] - d["ax"][1], d["ay"][0] - d["ay"][1],
                  d["az"][0] - d["az"][1]])
    w = v + a * dt
    np.testing.assert_allclose([mvx[0], mvy[0], mvz[0]], m1 * w, rtol=1e-14)
    np.testing.assert_allclose(k[0], m0 * m1 * (w @ w), rtol=1e-14)
    np.testing.assert_allclose(k[1], m1 * m0 * (w @ w), rtol=1e-14)
