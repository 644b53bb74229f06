"""Pinning of the oracle's post-Newtonian pair acceleration (oracle_pnacc).

  * Bit-exact against the golden vectors of the reference C backend
    (pnacc_kernel.c compiled from its own sources), FP64 and FP32, every
    order 0..8 at c = 128, order 7 at c = 16, and the j-prefix sweep.
  * Test-mass limit: with m_i = 0 the terms that survive are those of a test
    particle in the field of a mass moving at constant velocity, whose exact
    motion is the geodesic of the boosted Schwarzschild metric in harmonic
    coordinates.  The geodesic acceleration (Christoffel symbols by automatic
    differentiation) minus the Newtonian one and the PN series truncated at
    1PN, 2PN and 3PN must shrink as c^-4, c^-6 and c^-8: a wrong coefficient
    of any test-mass term (velocity-dependent ones included, through the
    boost) leaves a residual two powers of c larger.
  * 1PN energy: with the 1PN accelerations of both bodies, the
    Einstein-Infeld-Hoffmann energy is conserved up to O(c^-4); without them
    (or with any 1PN term wrong) dE/dt is O(c^-2).
  * 2.5PN: the relative acceleration a_1 - a_2 of a two-body system against
    the harmonic-gauge radiation-reaction acceleration (Iyer & Will 1995;
    Blanchet, Living Rev. Relativ. 17 (2014) 2, eq. 203), with the one
    coefficient in which the reference differs: its per-body v^2 v term is
    -1 (pn_terms.h:236) where the published one is -4/5, so the relative
    v^2 v coefficient is -2 eta M^2/r^3 instead of -8/5, and the circular-
    orbit energy loss is -34/5 eta^2 M^5/r^5 instead of Peters' -32/5.
"""
import numpy as np
import pytest

from oracle import Oracle, SPECS

torch = pytest.importorskip("torch")

FIELDS = ("mass", "rx", "ry", "rz", "eps2", "vx", "vy", "vz")
PREFIXES = (1, 2, 3, 7, 16, 33, 61)


def pn_scalars(order, c):
    inv1 = 1.0 / c
    return (order,) + tuple(inv1**k for k in range(1, 8))


def sets(g, tag, n=None):
    a = {k: g["%s_A_%s" % (tag, k)] for k in FIELDS}
    b = {k: g["%s_B_%s" % (tag, k)] for k in FIELDS}
    if n is not None:
        b = {k: v[:n] for k, v in b.items()}
    return a, b


@pytest.mark.parametrize("prec", ["float64", "float32"])
def test_pnacc_bitexact_vs_reference(golden, prec):
    g = golden("pairwise_%s.npz" % ("f64" if prec == "float64" else "f32"))
    orc = Oracle(prec)
    outs = SPECS["pnacc"][3]
    for soft in (0, 1):
        tag = "s%d" % soft
        a, b = sets(g, tag)
        for lab, ii, jj in (("AB", a, b), ("BA", b, a), ("AA", a, a)):
            res = orc.pairwise("pnacc", ii, jj, *pn_scalars(7, 128.0))
            for o, r in zip(outs, res):
                assert np.array_equal(r, g["%s_pnacc_%s_%s" % (tag, lab, o)])
        for k in PREFIXES:
            _, bk = sets(g, tag, k)
            res = orc.pairwise("pnacc", a, bk, *pn_scalars(7, 128.0))
            for o, r in zip(outs, res):
                assert np.array_equal(r, g["%s_pnacc_pre%d_%s" % (tag, k, o)])
        for lab, ii, jj in (("AB", a, b), ("AA", a, a)):
            for order in range(9):
                res = orc.pairwise("pnacc", ii, jj, *pn_scalars(order, 128.0))
                for o, r in zip(outs, res):
                    ref = g["%s_pnacc_o%d_%s_%s" % (tag, order, lab, o)]
                    assert np.array_equal(r, ref), (order, lab, o)
            res = orc.pairwise("pnacc", ii, jj, *pn_scalars(7, 16.0))
            for o, r in zip(outs, res):
                assert np.array_equal(r, g["%s_pnacc_c16_%s_%s" % (tag, lab,
                                                                   o)])


def test_pnacc_orders_are_distinct_and_low_orders_vanish(golden):
    g = golden("pairwise_f64.npz")
    r = {o: g["s0_pnacc_o%d_AB_pnax" % o] for o in range(9)}
    assert not np.any(r[0]) and not np.any(r[1])
    assert np.array_equal(r[2], r[3]) and np.array_equal(r[7], r[8])
    for lo, hi in ((3, 4), (4, 5), (5, 6), (6, 7)):
        assert not np.array_equal(r[lo], r[hi]), (lo, hi)


# ---------------------------------------------------------------------------
# Test-mass limit: geodesics of the boosted harmonic Schwarzschild metric.
# ---------------------------------------------------------------------------
def _boost(u, c):
    """Lambda = dX/dx from lab coordinates (ct, x) to the rest frame of a
    body moving with velocity u."""
    beta = torch.tensor(u, dtype=torch.float64) / c
    b2 = beta @ beta
    gam = 1.0 / torch.sqrt(1.0 - b2)
    L = torch.zeros(4, 4, dtype=torch.float64)
    L[0, 0] = gam
    L[0, 1:] = -gam * beta
    L[1:, 0] = -gam * beta
    L[1:, 1:] = torch.eye(3, dtype=torch.float64)
    if b2 > 0:
        L[1:, 1:] += (gam - 1.0) * torch.outer(beta, beta) / b2
    return L


def geodesic_acc(M, c, xj, u, xi, vi):
    """Coordinate acceleration d^2x/dt^2 of a test particle at xi with
    velocity vi, at lab time 0, in the field of mass M (G = 1) that is at xj
    at time 0 and moves with constant velocity u."""
    L = _boost(u, c)
    m = M / c**2
    xj4 = torch.tensor([0.0] + list(xj), dtype=torch.float64)
    eye = torch.eye(3, dtype=torch.float64)

    def metric(x4):
        X = L @ (x4 - xj4)
        Xs = X[1:]
        R = torch.sqrt(Xs @ Xs)
        N = Xs / R
        G = torch.zeros(4, 4, dtype=torch.float64)
        G[0, 0] = -(R - m) / (R + m)
        NN = torch.outer(N, N)
        G[1:, 1:] = (1 + m / R)**2 * (eye - NN) + (R + m) / (R - m) * NN
        return L.T @ G @ L

    x4 = torch.tensor([0.0] + list(xi), dtype=torch.float64)
    g = metric(x4)
    dg = torch.autograd.functional.jacobian(metric, x4)   # [mu, nu, alpha]
    ginv = torch.linalg.inv(g)
    # Gamma^mu_{ab} = 1/2 g^{mu nu} (d_a g_{nu b} + d_b g_{nu a} - d_nu g_ab)
    t = dg.permute(0, 2, 1) + dg - dg.permute(2, 0, 1)     # [nu, a, b]
    gam = 0.5 * torch.einsum("mn,nab->mab", ginv, t)
    w = torch.tensor([1.0] + [v / c for v in vi], dtype=torch.float64)
    A = torch.einsum("mab,a,b->m", gam, w, w)
    vi_t = torch.tensor(vi, dtype=torch.float64)
    return (-c**2 * (A[1:] - A[0] * vi_t / c)).numpy()


def _two(mi, xi, vi, mj, xj, vj):
    d = {"mass": np.array([mi, mj]), "eps2": np.zeros(2)}
    for k, (p, q) in zip(("rx", "ry", "rz"), zip(xi, xj)):
        d[k] = np.array([p, q])
    for k, (p, q) in zip(("vx", "vy", "vz"), zip(vi, vj)):
        d[k] = np.array([p, q])
    return d


@pytest.mark.parametrize("u", [(0.0, 0.0, 0.0), (0.3, -0.2, 0.25)])
def test_pnacc_test_mass_limit_matches_geodesic(u):
    orc = Oracle("float64")
    M = 1.0
    xj = (0.1, -0.2, 0.05)
    xi = (0.9, 0.5, -0.3)
    vi = (-0.3, 0.6, 0.2)
    d = _two(0.0, xi, vi, M, xj, u)
    r = np.array(xi) - np.array(xj)
    aN = -M * r / np.linalg.norm(r)**3
    res = {}
    for order in (0, 2, 4, 6):
        errs = []
        for c in (8.0, 16.0, 32.0):
            ex = geodesic_acc(M, c, xj, u, xi, vi)
            pn = orc.pairwise("pnacc", d, d, *pn_scalars(order, c))
            got = aN + np.array([pn[0][0], pn[1][0], pn[2][0]])
            errs.append(np.linalg.norm(got - ex))
        res[order] = errs
    # the truncation error falls by 2^(order + 2) per doubling of c
    for order, errs in res.items():
        p = order + 2
        for e0, e1 in zip(errs, errs[1:]):
            ratio = e0 / e1
            assert 0.6 * 2**p < ratio < 1.6 * 2**p, (order, errs)
    # 3.5PN and 2.5PN carry m_i: nothing changes for a test mass
    for c in (8.0, 32.0):
        p6 = orc.pairwise("pnacc", d, d, *pn_scalars(6, c))
        p7 = orc.pairwise("pnacc", d, d, *pn_scalars(7, c))
        assert all(np.array_equal(x, y) for x, y in zip(p6, p7))


# ---------------------------------------------------------------------------
# 1PN energy conservation (Einstein-Infeld-Hoffmann).
# ---------------------------------------------------------------------------
def eih_energy(m, x, v, c):
    """Two-body 1PN energy (G = 1): Newtonian terms plus
    1/c^2 [3/8 sum m v^4 + m1 m2/(2r) (3 v1^2 + 3 v2^2 - 7 v1.v2
    - (n.v1)(n.v2)) + m1 m2 (m1 + m2)/(2 r^2)]."""
    r12 = x[0] - x[1]
    r = torch.sqrt(r12 @ r12)
    n = r12 / r
    v1, v2 = v[0], v[1]
    m1, m2 = m
    e = 0.5 * m1 * (v1 @ v1) + 0.5 * m2 * (v2 @ v2) - m1 * m2 / r
    e1 = (3.0 / 8.0) * (m1 * (v1 @ v1)**2 + m2 * (v2 @ v2)**2)
    e1 = e1 + m1 * m2 / (2 * r) * (3 * (v1 @ v1) + 3 * (v2 @ v2)
                                   - 7 * (v1 @ v2) - (n @ v1) * (n @ v2))
    e1 = e1 + m1 * m2 * (m1 + m2) / (2 * r**2)
    return e + e1 / c**2


def edot(m, x, v, a, c):
    xt = torch.tensor(x, dtype=torch.float64, requires_grad=True)
    vt = torch.tensor(v, dtype=torch.float64, requires_grad=True)
    E = eih_energy(m, xt, vt, c)
    gx, gv = torch.autograd.grad(E, (xt, vt))
    return float((gx * vt.detach()).sum() + (gv * torch.tensor(a)).sum())


def test_pnacc_1pn_conserves_eih_energy():
    orc = Oracle("float64")
    m = (0.7, 0.4)
    x = np.array([[0.3, 0.1, -0.2], [-0.5, 0.2, 0.4]])
    v = np.array([[0.2, 0.5, -0.1], [-0.3, -0.6, 0.2]])
    d = _two(m[0], x[0], v[0], m[1], x[1], v[1])
    r12 = x[0] - x[1]
    r = np.linalg.norm(r12)
    aN = np.array([-m[1] * r12 / r**3, m[0] * r12 / r**3])
    with_pn, without = [], []
    for c in (20.0, 40.0, 80.0):
        pn = orc.pairwise("pnacc", d, d, *pn_scalars(2, c))
        a = aN + np.stack(pn, axis=1)
        with_pn.append(abs(edot(m, x, v, a, c)))
        without.append(abs(edot(m, x, v, aN, c)))
    for e0, e1 in zip(with_pn, with_pn[1:]):
        assert 10 < e0 / e1 < 25, with_pn          # O(c^-4)
    for e0, e1 in zip(without, without[1:]):
        assert 3 < e0 / e1 < 5, without            # O(c^-2)
    assert with_pn[-1] < 1e-2 * without[-1]


# ---------------------------------------------------------------------------
# 2.5PN radiation reaction.
# ---------------------------------------------------------------------------
def pn25_pair(orc, m1, m2, x1, x2, v1, v2, c):
    """Per-body 2.5PN accelerations: order-5 minus order-4 outputs."""
    d = _two(m1, x1, v1, m2, x2, v2)
    p5 = np.stack(orc.pairwise("pnacc", d, d, *pn_scalars(5, c)), axis=1)
    p4 = np.stack(orc.pairwise("pnacc", d, d, *pn_scalars(4, c)), axis=1)
    return p5 - p4


@pytest.mark.parametrize("masses", [(0.5, 0.5), (0.8, 0.2), (0.3, 1.1)])
def test_pnacc_25pn_relative_acceleration(masses):
    orc = Oracle("float64")
    m1, m2 = masses
    M = m1 + m2
    eta = m1 * m2 / M**2
    c = 10.0
    x = np.array([0.6, -0.2, 0.3])
    v = np.array([0.3, 0.7, -0.2])
    # centre-of-mass frame
    x1, x2 = m2 / M * x, -m1 / M * x
    v1, v2 = m2 / M * v, -m1 / M * v
    a = pn25_pair(orc, m1, m2, x1, x2, v1, v2, c)
    arel = a[0] - a[1]
    r = np.linalg.norm(x)
    n = x / r
    rdot = n @ v
    v2s = v @ v
    k = eta * M**2 / r**3 / c**5
    # Iyer-Will harmonic gauge: 8/5 [(3 v^2 + 17/3 M/r) rdot n
    # - (v^2 + 3 M/r) v]; the reference's per-body -v^2 v term makes the
    # v^2 v coefficient -2 instead of -8/5
    expect = k * ((24.0 / 5 * v2s + 136.0 / 15 * M / r) * rdot * n
                  - (2.0 * v2s + 24.0 / 5 * M / r) * v)
    published = k * ((24.0 / 5 * v2s + 136.0 / 15 * M / r) * rdot * n
                     - (8.0 / 5 * v2s + 24.0 / 5 * M / r) * v)
    np.testing.assert_allclose(arel, expect, rtol=1e-9, atol=0)
    assert np.linalg.norm(arel - published) > 1e-3 * np.linalg.norm(arel)


def test_pnacc_25pn_circular_energy_loss():
    """dE_N/dt = sum m v.a_2.5PN on a circular orbit: -34/5 eta^2 M^5/r^5
    with the reference's coefficient (Peters' quadrupole loss is -32/5)."""
    orc = Oracle("float64")
    m1, m2 = 0.7, 0.3
    M = m1 + m2
    eta = m1 * m2 / M**2
    c = 10.0
    r = 1.3
    x = np.array([r, 0.0, 0.0])
    v = np.array([0.0, np.sqrt(M / r), 0.0])
    x1, x2 = m2 / M * x, -m1 / M * x
    v1, v2 = m2 / M * v, -m1 / M * v
    a = pn25_pair(orc, m1, m2, x1, x2, v1, v2, c)
    de = m1 * (v1 @ a[0]) + m2 * (v2 @ a[1])
    peters = -32.0 / 5 * eta**2 * M**5 / r**5 / c**5
    np.testing.assert_allclose(de, 34.0 / 32.0 * peters, rtol=1e-9)
