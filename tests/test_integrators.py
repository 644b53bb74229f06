"""Host logic of the drivers on the CPU: block time steps, Hermite order of
convergence, SIA hierarchy, Sakura, the Plummer generator and the extension
layer's call pattern.  Forces come from the test-only oracle backend
(tests/oracle_backend.py) on CPU-resident particle systems."""
import math

import numpy as np
import pytest
import torch

from oracle import Oracle
from oracle_backend import OracleKernel, install_hip, install_oracle


@pytest.fixture(autouse=True)
def oracle_forces():
    install_oracle("float64")
    yield
    install_hip("float64")
    # PN runs switch the class-wide flag on, as the reference does
    from tupan_amd.particles import ParticleSystem
    ParticleSystem.include_pn_corrections = False


def make_ps(d):
    from tupan_amd.particles import ParticleSystem
    return ParticleSystem.from_numpy(d, device="cpu")


def kepler_pair(e=0.0):
    """Two equal masses (total 1) on a circular (e=0) or eccentric orbit,
    centre of mass at rest at the origin."""
    r = 1.0
    v = math.sqrt((1 + e) / r)
    d = {"mass": np.array([0.5, 0.5]), "eps2": np.zeros(2),
         "rx": np.array([0.5, -0.5]), "ry": np.zeros(2), "rz": np.zeros(2),
         "vx": np.zeros(2), "vy": np.array([v / 2, -v / 2]),
         "vz": np.zeros(2)}
    return d


def test_min_block_tstep():
    from tupan_amd.integrator.hermite import Hermite
    ps = make_ps(kepler_pair())
    it = Hermite(0.01, 0.0, ps, "ahermite4")
    ps.tstep[...] = 0.01
    type(ps).t_curr = 0.0
    assert it.get_min_block_tstep(ps, 1.0) == 2.0**-7
    type(ps).t_curr = 2.0**-8
    assert it.get_min_block_tstep(ps, 1.0) == 2.0**-8
    type(ps).t_curr = 0.0
    assert it.get_min_block_tstep(ps, 2.0**-9) == 2.0**-9
    assert it.get_min_block_tstep(ps, -1.0) == -(2.0**-7)
    type(ps).t_curr = 0.0


def test_base_tstep():
    from tupan_amd.integrator.hermite import Hermite
    ps = make_ps(kepler_pair())
    it = Hermite(0.125, 0.0, ps, "hermite4")
    assert it.get_base_tstep(1.0) == 0.125
    type(ps).t_curr = 0.95
    assert abs(it.get_base_tstep(1.0) - 0.05) < 1e-15
    type(ps).t_curr = 0.0


def run_fixed(method, dt, t_end, d):
    from tupan_amd.integrator import Integrator
    ps = make_ps(d)
    it = Integrator(dt, 0.0, ps, method=method)
    while it.time < t_end - 1e-12:
        it.evolve_step(t_end)
    return it.particle_system


@pytest.mark.parametrize("method,order", [("hermite2", 2), ("hermite4", 4),
                                          ("hermite6", 6), ("hermite8", 8)])
def test_hermite_order_of_convergence(method, order):
    """Position error after t = 1 on an eccentric two-body orbit against
    the exact Kepler solution: halving the step divides it by ~2^order."""
    d = kepler_pair(0.3)
    orc = Oracle("float64")
    exact = orc.kepler(d, 1.0)
    errs = []
    steps = (1 / 16, 1 / 32) if order < 6 else (1 / 8, 1 / 16)
    for dt in steps:
        ps = run_fixed(method, dt, 1.0, d)
        err = abs(float(ps.rx[0]) - exact[0][0]) + \
            abs(float(ps.ry[0]) - exact[1][0])
        errs.append(err)
    ratio = errs[0] / errs[1]
    assert ratio > 2**order * 0.6, (errs, ratio)


def test_sakura_two_body_is_exact():
    """Sakura advances each pair exactly: a lone pair follows the Kepler
    orbit to solver precision whatever the step."""
    d = kepler_pair(0.5)
    ps = run_fixed("sakura", 0.25, 1.0, d)
    exact = Oracle("float64").kepler(d, 1.0)
    assert abs(float(ps.rx[0]) - exact[0][0]) < 1e-11
    assert abs(float(ps.vy[1]) - exact[4][1]) < 1e-11


def test_sia_leaf_is_kepler_and_hierarchy_aligns():
    """sia21h: a pair ends in the FewBody Kepler leaf (exact), and a larger
    system ends every particle at the common end time."""
    d = kepler_pair(0.2)
    ps = run_fixed("sia21h.dkd", 0.125, 0.5, d)
    exact = Oracle("float64").kepler(d, 0.5)
    assert abs(float(ps.rx[0]) - exact[0][0]) < 1e-11
    from tupan_amd.ics.plummer import make_plummer
    p = make_plummer(48, 0.05, ("equalmass",), seed=3, device="cpu")
    from tupan_amd.integrator import Integrator
    type(p).t_curr = 0.0
    it = Integrator(0.05, 0.0, p, method="sia21h.dkd")
    e0 = p.kinetic_energy + p.potential_energy
    it.evolve_step(0.25)
    q = it.particle_system
    steps = q.tstep.numpy()
    # every level's step is the base step (eta) over a power of two
    lg = np.log2(np.abs(steps) / 0.05)
    assert np.allclose(lg, np.round(lg))
    assert q.n == 48 and sorted(q.id.tolist()) == list(range(48))
    e1 = q.kinetic_energy + q.potential_energy
    assert abs((e1 - e0) / e0) < 1e-3


@pytest.mark.parametrize("method", ["sia21h.dkd", "sia22h.kdk",
                                    "sia22a.dkd", "sia44s.kdk"])
def test_sia_inplace_partition_equals_copy_join(method):
    """The in-place [slow | fast] partition gives the particle order and
    every value of the reference's ps[mask] copies + append, bit for bit,
    over several steps of a hierarchy several levels deep."""
    from tupan_amd.ics.imf import IMF
    from tupan_amd.ics.plummer import make_plummer
    from tupan_amd.integrator import sia
    runs = []
    for inplace in (False, True):
        p = make_plummer(40, 0.02, IMF.power_law(-2.35, 0.1, 10.0), seed=7,
                         device="cpu")
        type(p).t_curr = 0.0
        it = sia.SIA(0.05, 0.0, p, method, inplace=inplace)
        for k in range(3):
            it.evolve_step(0.05 * (k + 1))
        runs.append((it.ps, type(p).t_curr))
    (a, ta), (b, tb) = runs
    assert ta == tb
    for k in ("id", "mass", "rx", "ry", "rz", "vx", "vy", "vz", "time",
              "nstep", "tstep"):
        assert torch.equal(getattr(a, k), getattr(b, k)), k
    if method.startswith("sia21h"):
        # the hierarchy went several levels deep and reordered particles
        assert len(set(a.tstep.tolist())) >= 3
        assert not torch.equal(a.id, torch.arange(a.n))


def test_plummer_invariants_and_rng_sequence():
    from tupan_amd.ics.plummer import make_plummer
    n = 64
    eps = 0.01
    ps = make_plummer(n, eps, ("equalmass",), seed=1, device="cpu")
    assert abs(ps.total_mass - 1.0) < 1e-14
    np.testing.assert_allclose(ps.com_r, 0, atol=1e-15)
    np.testing.assert_allclose(ps.com_v, 0, atol=1e-15)
    ke, pe = ps.kinetic_energy, ps.potential_energy
    assert abs(2 * ke / -pe - 1) < 1e-12
    # eps2 is not rescaled with the positions (plummer.py:131,
    # body.py:449-461), so r_vir = 1 only up to the softening's effect
    assert abs(ps.virial_radius - 1) < 1e-3
    np.testing.assert_array_equal(ps.eps2.numpy(), np.full(n, eps**2 / 2))
    np.testing.assert_array_equal(ps.mass.numpy(), np.full(n, 1.0 / n))
    hard = make_plummer(n, 0.0, ("equalmass",), seed=1, device="cpu")
    assert abs(hard.virial_radius - 1) < 1e-12


def test_von_neumann_batches_equal_scalar_loop():
    """The batched rejection sampler returns the reference's scalar loop
    values (plummer.py:77-86) and leaves the generator where that loop
    does."""
    from tupan_amd.ics.plummer import _von_neumann
    for seed, n in ((1, 5), (2, 100), (3, 1000)):
        a = np.random.RandomState(seed)
        b = np.random.RandomState(seed)
        got = _von_neumann(a, n)
        ref = []
        while len(ref) < n:
            r1 = b.random_sample()
            r2 = b.random_sample()
            if r2 < r1:
                ref.append(r2)
        np.testing.assert_array_equal(got, ref)
        assert a.random_sample() == b.random_sample()


def test_plummer_positions_follow_reference_sequence():
    """Positions are the reference's draws after the state restore: a
    permutation, random(n), uniform(-1,1,n), uniform(0,2pi,n), scaled by the
    virial rescaling (plummer.py:65-75)."""
    from tupan_amd.ics.plummer import make_plummer
    n = 32
    ps = make_plummer(n, 0.0, ("equalmass",), seed=5, device="cpu")
    rs = np.random.RandomState(5)
    irand = rs.permutation(np.arange(n))
    mrand = (irand + rs.random_sample(n)) * 0.999 / n
    radius = 1.0 / np.sqrt(np.power(mrand, -2.0 / 3.0) - 1.0)
    theta = np.arccos(rs.uniform(-1.0, 1.0, size=n))
    phi = rs.uniform(0.0, 2.0 * np.pi, size=n)
    rz = radius * np.cos(theta)
    rx = radius * np.sin(theta) * np.cos(phi)
    # the model is shifted to its centre of mass and scaled by one factor
    got = ps.rz.numpy()
    gx = ps.rx.numpy()
    ref = rz - rz.mean()
    refx = rx - rx.mean()
    s = np.dot(got, ref) / np.dot(ref, ref)
    np.testing.assert_allclose(got, s * ref, rtol=1e-10, atol=1e-12)
    np.testing.assert_allclose(gx, s * refx, rtol=1e-10, atol=1e-12)


def test_extension_call_pattern_hermite4():
    """ahermite4 does 1 Tstep and 3 AccJerk all-pairs calls per step
    (hermite.py:80,97,347,396)."""
    from tupan_amd.ics.plummer import make_plummer
    from tupan_amd.integrator import Integrator
    p = make_plummer(16, 0.01, ("equalmass",), seed=2, device="cpu")
    type(p).t_curr = 0.0
    it = Integrator(0.01, 0.0, p, method="ahermite4")
    it.evolve_step(1.0)
    OracleKernel.calls.clear()
    it.evolve_step(1.0)
    names = [c[0] for c in OracleKernel.calls]
    assert names.count("acc_jerk") == 3 and names.count("tstep") == 1


def test_particle_system_structure():
    from tupan_amd.particles import ParticleSystem
    ps = ParticleSystem(3, 2, device="cpu")
    assert ps.n == 5 and set(ps.members) == {"bodies", "stars"}
    ps.rx[...] = torch.arange(5, dtype=torch.float64)
    assert ps.stars.rx.tolist() == [3.0, 4.0]
    ps.stars.rx[...] = 9.0
    assert ps.rx.tolist() == [0, 1, 2, 9, 9]
    sub = ps[torch.tensor([True, False, True, False, True])]
    assert sub.n == 3 and sub.rx.tolist() == [0, 2, 9]
    sub.rx[...] = -1
    assert ps.rx.tolist() == [0, 1, 2, 9, 9]
    c = ps.copy()
    c.append(sub)
    assert c.n == 8 and c.rx.tolist()[-3:] == [-1, -1, -1]
    ps.register_auxiliary_attribute("phi", "real")
    with pytest.raises(ValueError):
        ps.register_auxiliary_attribute("phi", "real")
    st = ps.get_state()
    assert st.dtype.names[:3] == ("id", "mass", "eps2")


@pytest.mark.parametrize("method", ["nreg", "anreg"])
def test_nreg_two_body_conserves_energy(method):
    """NREG's time-transformed leapfrog is exact in shape for a Kepler pair
    (the logarithmic-Hamiltonian leapfrog): after many steps of an
    eccentric orbit the energy and angular momentum stay at round-off and
    the separation stays on the orbit's [pericentre, apocentre]."""
    d = kepler_pair(0.6)
    ps = make_ps(d)
    from tupan_amd.integrator import Integrator
    type(ps).t_curr = 0.0
    it = Integrator(0.1, 0.0, ps, method=method)
    e0 = ps.kinetic_energy + ps.potential_energy
    L0 = float(ps.mass[0] * (ps.rx[0] * ps.vy[0] - ps.ry[0] * ps.vx[0])
               + ps.mass[1] * (ps.rx[1] * ps.vy[1] - ps.ry[1] * ps.vx[1]))
    seps = []
    for k in range(200):
        it.evolve_step(1e9)
        q = it.particle_system
        seps.append(float(((q.rx[0] - q.rx[1])**2
                           + (q.ry[0] - q.ry[1])**2)**0.5))
    q = it.particle_system
    e1 = q.kinetic_energy + q.potential_energy
    L1 = float(q.mass[0] * (q.rx[0] * q.vy[0] - q.ry[0] * q.vx[0])
               + q.mass[1] * (q.rx[1] * q.vy[1] - q.ry[1] * q.vx[1]))
    assert abs((e1 - e0) / e0) < 1e-12
    assert abs((L1 - L0) / L0) < 1e-12
    # orbit: a = 1 / (2 |E| / m) with total mass 1, e = 0.6 -> r in [a(1-e), a(1+e)]
    a = 1.0 / (-2 * e0 / (0.5 * 0.5))
    assert min(seps) > a * 0.4 * (1 - 1e-9) and max(seps) < a * 1.6 * (1 + 1e-9)
    assert it.time > 0


def test_nreg_plummer_second_order():
    """On a 24-body Plummer sphere the energy error over a fixed span of
    regularised time drops about fourfold when the step is halved."""
    from tupan_amd.ics.plummer import make_plummer
    from tupan_amd.integrator import Integrator
    errs = []
    for h, n in ((0.02, 10), (0.01, 20)):
        p = make_plummer(24, 0.02, ("equalmass",), seed=5, device="cpu")
        type(p).t_curr = 0.0
        e0 = p.kinetic_energy + p.potential_energy
        it = Integrator(h, 0.0, p, method="nreg")
        for _ in range(n):
            it.evolve_step(1e9)
        q = it.particle_system
        errs.append(abs((q.kinetic_energy + q.potential_energy - e0) / e0))
    assert errs[0] < 1e-3
    assert 2.5 < errs[0] / errs[1] < 6.5, errs


def pn_group(seed=3):
    """Six bodies: a compact binary inside a small bound group, speeds of a
    few tenths, so that (v/c)^2 ~ 1e-3 at c = 20."""
    rng = np.random.default_rng(seed)
    n = 6
    d = {"mass": np.full(n, 1.0 / n), "eps2": np.zeros(n)}
    pos = rng.normal(0, 0.5, (n, 3))
    vel = rng.normal(0, 0.3, (n, 3))
    pos[1] = pos[0] + np.array([0.05, 0.0, 0.0])
    vel[1] = vel[0] + np.array([0.0, np.sqrt(2 / n / 0.05), 0.0])
    pos -= (d["mass"][:, None] * pos).sum(0)
    vel -= (d["mass"][:, None] * vel).sum(0)
    for k, (p, v) in enumerate(zip(("rx", "ry", "rz"), ("vx", "vy", "vz"))):
        d[p] = pos[:, k].copy()
        d[v] = vel[:, k].copy()
    return d


def pn_run(method, tau, nsteps, order=7, c=20.0):
    from tupan_amd.integrator import Integrator
    p = make_ps(pn_group())
    type(p).t_curr = 0.0
    it = Integrator(tau, 0.0, p, method=method, pn_order=order, clight=c)
    ps = it.particle_system
    e0 = ps.kinetic_energy + ps.potential_energy
    p0, l0, c0 = ps.linear_momentum, ps.angular_momentum, ps.com_r
    for k in range(nsteps):
        it.evolve_step(tau * (k + 1))
    q = it.particle_system
    return q, e0, p0, l0, c0, type(p).t_curr


@pytest.mark.parametrize("method", ["sia21s.dkd", "sia22a.kdk",
                                    "sia44s.dkd"])
def test_sia_pn_bookkeeping_conserves_momenta(method):
    """With the PN impulse and torque kept in pn_mv and pn_am
    (body.py:496-528), the corrected linear and angular momenta are
    conserved to rounding, and the corrected centre of mass moves with the
    corrected momentum, whatever the PN forces do; PN terms did act."""
    q, e0, p0, l0, c0, t = pn_run(method, 1.0 / 64, 8)
    assert q.include_pn_corrections
    assert float(q.pn_ke.abs().max()) > 0 and float(q.pn_mvx.abs().max()) > 0
    assert np.max(np.abs(q.linear_momentum - p0)) < 1e-15
    assert np.max(np.abs(q.angular_momentum - l0)) < 1e-14
    np.testing.assert_allclose(q.com_r, c0 + p0 * t / q.total_mass,
                               rtol=0, atol=1e-15)
    e1 = q.kinetic_energy + q.potential_energy
    assert abs((e1 - e0) / e0) < 5e-3   # 8 steps per binary orbit


def test_sia_pn_energy_second_order():
    """sia21s.dkd with PN kicks: the PN-corrected energy error drops about
    fourfold when the shared step is halved."""
    errs = []
    for tau, n in ((1.0 / 128, 16), (1.0 / 256, 32)):
        q, e0, *_ = pn_run("sia21s.dkd", tau, n)
        errs.append(abs((q.kinetic_energy + q.potential_energy - e0) / e0))
    assert 3.0 < errs[0] / errs[1] < 5.5, errs


def test_sia_pn_order_zero_is_newtonian():
    """pn_order 0 leaves the Newtonian integration untouched; a PN run
    departs from it at O(1/c^2)."""
    from tupan_amd.integrator import Integrator
    out = []
    for kw in ({}, {"pn_order": 7, "clight": 80.0},
               {"pn_order": 7, "clight": 160.0}):
        p = make_ps(pn_group())
        type(p).t_curr = 0.0
        it = Integrator(1.0 / 64, 0.0, p, method="sia21s.dkd", **kw)
        for k in range(4):
            it.evolve_step((k + 1) / 64)
        q = it.particle_system
        assert q.include_pn_corrections == bool(kw)
        assert ("pn_ke" in q.__dict__) == bool(kw)
        out.append(q)
    newt, pn80, pn160 = out
    d80 = float((newt.vx - pn80.vx).abs().max())
    d160 = float((newt.vx - pn160.vx).abs().max())
    assert 3.0 < d80 / d160 < 5.0, (d80, d160)
    with pytest.raises(TypeError):
        Integrator(0.1, 0.0, make_ps(pn_group()), method="sia21s.dkd",
                   pn_order=2)


def test_sia_pn_pair_leaf_raises_as_reference():
    """A hierarchical step that isolates the binary hands FewBody a pair,
    which the reference refuses with PN corrections (fewbody.py:34-40)."""
    with pytest.raises(NotImplementedError):
        pn_run("sia21h.dkd", 1.0 / 16, 2)


def test_diagnostic_report_file(tmp_path):
    """Diagnostic(fname, time, ps, report_freq) writes the reference's
    header line, then one line per report_freq reports with 13 columns;
    te = ke + pe and eerr = (te - te0)/(-pe) as printed."""
    from tupan_amd.diagnostics import Diagnostic
    from tupan_amd.integrator import Integrator
    from tupan_amd.ics.plummer import make_plummer
    p = make_plummer(32, 0.05, ("equalmass",), seed=2, device="cpu")
    type(p).t_curr = 0.0
    fname = str(tmp_path / "diag.txt")
    rep = Diagnostic(fname, 0.0, p, report_freq=2)
    it = Integrator(1.0 / 64, 0.0, p, method="ahermite4", reporter=rep)
    for k in range(6):
        it.evolve_step((k + 1) / 64)
    lines = open(fname).read().splitlines()
    assert lines[0].split() == list(Diagnostic.HEADER)
    assert lines[0].startswith("#00:time      #01:dtime  #02:ke     ")
    body = [[float(x) for x in ln.split()] for ln in lines[1:]]
    assert len(body) == 4 and all(len(r) == 13 for r in body)   # 7 reports
    for r in body:
        t, dt, ke, pe, te, vir, eerr = r[:7]
        assert abs(te - (ke + pe)) < 1e-3 * abs(te)   # ke, pe: 4 digits
        assert abs(vir - (2 * ke + pe)) < 1e-2
    assert body[0][0] == 0.0 and body[0][6] == 0.0
    assert rep.rows[-1]["time"] == it.time


def test_sakura_device_steps_argument_checks():
    """Device Sakura steps: evolve needs device_steps and at least one step
    per graph, and on a host particle system (or the oracle backend) the
    device path refuses to start."""
    from tupan_amd.integrator import Integrator
    it = Integrator(1.0 / 64, 0.0, make_ps(kepler_pair()), method="asakura",
                    device_steps=True)
    with pytest.raises(RuntimeError):
        it.evolve_step(1.0)
    it = Integrator(1.0 / 64, 0.0, make_ps(kepler_pair()), method="sakura")
    with pytest.raises(RuntimeError):
        it.integrator.evolve(1.0)
    it = Integrator(1.0 / 64, 0.0, make_ps(kepler_pair()), method="sakura",
                    device_steps=True)
    with pytest.raises(ValueError):
        it.integrator.evolve(1.0, steps_per_graph=0)
    with pytest.raises(RuntimeError):
        it.evolve_step(1.0)


def test_hermite_device_steps_argument_checks():
    """Device Hermite steps take the same checks as the Sakura ones: evolve
    needs device_steps and at least one step per graph, and on a host
    particle system the device path refuses to start."""
    from tupan_amd.integrator import Integrator
    it = Integrator(1.0 / 64, 0.0, make_ps(kepler_pair()), method="ahermite4")
    with pytest.raises(RuntimeError):
        it.integrator.evolve(1.0)
    it = Integrator(1.0 / 64, 0.0, make_ps(kepler_pair()), method="ahermite4",
                    device_steps=True)
    with pytest.raises(ValueError):
        it.integrator.evolve(1.0, steps_per_graph=0)
    with pytest.raises(RuntimeError):
        it.evolve_step(1.0)


def test_sia_device_steps_argument_checks():
    """Device SIA steps exist for the shared-step 's' and 'a' variants:
    evolve needs device_steps and at least one step per graph, a
    hierarchical method or a host particle system refuses to start."""
    from tupan_amd.integrator import Integrator
    it = Integrator(1.0 / 64, 0.0, make_ps(kepler_pair()), method="sia21a.dkd")
    with pytest.raises(RuntimeError):
        it.integrator.evolve(1.0)
    it = Integrator(1.0 / 64, 0.0, make_ps(kepler_pair()), method="sia21a.dkd",
                    device_steps=True)
    with pytest.raises(ValueError):
        it.integrator.evolve(1.0, steps_per_graph=0)
    with pytest.raises(RuntimeError):
        it.evolve_step(1.0)
    it = Integrator(1.0 / 64, 0.0, make_ps(kepler_pair()), method="sia21h.dkd",
                    device_steps=True)
    with pytest.raises(RuntimeError):
        it.evolve_step(1.0)
