"""Ensembles of independent systems: one launch over all systems against the
per-system oracle, across lane-packing policies, and the ensemble driver
against the single-system Sakura driver.

Bar for dr, dv against the oracle: 1e-9 of each particle's vector norm.  The
device transcendentals differ from glibc in the last bit; the Laguerre
iteration stops at |ds| <= 2^-42, so converged positions agree to ~1e-13 of
|r|, and dr = mu (r1 - r0) is a difference that amplifies that by |r|/|dr|
(up to ~1e3 on these random systems at dt = 0.3)."""
import numpy as np
import pytest

from conftest import vec_rel_err

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

F8 = ("mass", "rx", "ry", "rz", "eps2", "vx", "vy", "vz")


def random_ensemble(nsys, seed, soft=False):
    rng = np.random.default_rng(seed)
    sizes = rng.integers(1, 65, nsys)
    sizes[:3] = (1, 2, 64)
    off = np.zeros(nsys + 1, np.int64)
    off[1:] = np.cumsum(sizes)
    n = int(off[-1])
    d = {k: rng.normal(0, 1, n) for k in ("rx", "ry", "rz")}
    d.update({k: rng.normal(0, 0.6, n) for k in ("vx", "vy", "vz")})
    d["mass"] = rng.random(n) * 0.1 + 0.01
    d["eps2"] = rng.random(n) * 1e-6 if soft else np.zeros(n)
    return d, off


@pytest.mark.parametrize("min_lanes", [8, 16, 32, 64])
def test_ensemble_sakura_vs_oracle(min_lanes):
    from oracle import Oracle
    from tupan_amd.lib.ensemble import Ensemble
    from tupan_amd.particles import ParticleSystem
    d, off = random_ensemble(37, 3, soft=(min_lanes == 16))
    ps = ParticleSystem.from_numpy(d, device="cuda:0")
    ens = Ensemble(off.astype(np.int32), min_lanes=min_lanes)
    orc = Oracle("float64")
    for flag, dt in ((-1, 1.0 / 64), (1, 0.3), (2, 0.1), (-2, 0.05)):
        got = [t.cpu().numpy().copy() for t in ens.sakura(ps, dt, flag)]
        for s in range(len(off) - 1):
            a, b = off[s], off[s + 1]
            sub = {k: v[a:b] for k, v in d.items()}
            ref = orc.sakura(sub, sub, dt, flag)
            assert vec_rel_err([g[a:b] for g in got[:3]], ref[:3]) <= 1e-9
            assert vec_rel_err([g[a:b] for g in got[3:]], ref[3:]) <= 1e-9


def test_ensemble_lane_policies_bit_identical():
    from tupan_amd.lib.ensemble import Ensemble
    from tupan_amd.particles import ParticleSystem
    d, off = random_ensemble(50, 9)
    ps = ParticleSystem.from_numpy(d, device="cuda:0")
    outs = []
    for L in (8, 16, 32, 64):
        ens = Ensemble(off.astype(np.int32), min_lanes=L)
        outs.append([t.cpu().numpy().copy() for t in ens.sakura(ps, 0.02,
                                                                -1)])
    for o in outs[1:]:
        for x, y in zip(o, outs[0]):
            assert np.array_equal(x, y)


def test_ensemble_phi_vs_oracle():
    from oracle import Oracle
    from tupan_amd.lib.ensemble import Ensemble
    from tupan_amd.particles import ParticleSystem
    d, off = random_ensemble(29, 5, soft=True)
    ps = ParticleSystem.from_numpy(d, device="cuda:0")
    ens = Ensemble(off.astype(np.int32))
    phi = ens.phi(ps).cpu().numpy()
    orc = Oracle("float64")
    for s in range(len(off) - 1):
        a, b = off[s], off[s + 1]
        sub = {k: v[a:b] for k, v in d.items()}
        (ref,) = orc.phi(sub, sub)
        assert np.array_equal(phi[a:b], ref)  # reference order, IEEE ops


def test_make_ensemble_matches_make_plummer():
    from tupan_amd.ics.ensemble import make_ensemble, system_seed, \
        system_sizes
    from tupan_amd.ics.imf import IMF
    from tupan_amd.ics.plummer import make_plummer
    nsys = 6
    ps, ens = make_ensemble(nsys=nsys, nmin=3, nmax=20, seed=4,
                            device="cuda:0")
    sizes = system_sizes(nsys, 3, 20, 4)
    off = np.concatenate([[0], np.cumsum(sizes)])
    for s in range(nsys):
        one = make_plummer(int(sizes[s]), 0.0, IMF.log_uniform(0.1, 1.0),
                           seed=system_seed(4, s), device="cuda:0")
        for k in ("mass", "rx", "vz"):
            np.testing.assert_allclose(
                getattr(ps, k)[off[s]:off[s + 1]].cpu().numpy(),
                getattr(one, k).cpu().numpy(), rtol=1e-11, atol=1e-13)
    ke, pe = ens.energies(ps)
    ratio = (2 * ke / -pe).cpu().numpy()
    np.testing.assert_allclose(ratio, 1.0, rtol=1e-12)


def test_ensemble_driver_vs_single_system_sakura():
    from oracle_backend import install_hip
    from tupan_amd.ics.ensemble import make_ensemble
    from tupan_amd.integrator import Integrator
    from tupan_amd.integrator.ensemble import SakuraEnsemble
    from tupan_amd.particles import ParticleSystem
    install_hip("float64")
    ps, ens = make_ensemble(nsys=5, nmin=3, nmax=12, seed=2, device="cuda:0")
    off = ens.offsets
    host = ps.to_numpy()
    drv = SakuraEnsemble(ps, ens)
    for _ in range(3):
        drv.step(1.0 / 64)
    for s in range(5):
        a, b = off[s], off[s + 1]
        one = ParticleSystem.from_numpy({k: v[a:b] for k, v in host.items()},
                                        device="cuda:0")
        type(one).t_curr = 0.0
        it = Integrator(1.0 / 64, 0.0, one, method="sakura")
        for _ in range(3):
            it.evolve_step(1.0)
        q = it.particle_system
        for k in ("rx", "ry", "rz", "vx", "vy", "vz"):
            np.testing.assert_allclose(getattr(ps, k)[a:b].cpu().numpy(),
                                       getattr(q, k).cpu().numpy(),
                                       rtol=0, atol=1e-12)


def test_ensemble_graph_replay_equals_eager():
    """A captured block of Sakura ensemble steps, replayed several times with
    eager work and synchronisations in between, gives the eager steps' numbers
    after every replay; evolve_graph equals evolve."""
    from tupan_amd.ics.ensemble import make_ensemble
    from tupan_amd.integrator.ensemble import SakuraEnsemble
    from tupan_amd.lib.ensemble import Ensemble
    from tupan_amd.particles import ParticleSystem
    ps1, ens1 = make_ensemble(nsys=64, nmin=3, nmax=40, seed=3,
                              device="cuda:0")
    host = ps1.to_numpy()
    ps2 = ParticleSystem.from_numpy(host, device="cuda:0")
    ens2 = Ensemble(ens1.offsets)
    eager = SakuraEnsemble(ps1, ens1)
    graphed = SakuraEnsemble(ps2, ens2)
    tau = 1.0 / 64
    g = graphed.capture(tau, 4)
    keys = ("rx", "ry", "rz", "vx", "vy", "vz")
    for rep in range(3):
        g.replay()
        for _ in range(4):
            eager.step(tau)
        phi = ens1.phi(ps1).sum().item()  # eager work between replays
        torch.cuda.synchronize()
        assert phi == phi
        for k in keys:
            assert torch.equal(getattr(ps1, k), getattr(ps2, k)), (rep, k)
    ens2.check_status()

    ps3 = ParticleSystem.from_numpy(host, device="cuda:0")
    ps4 = ParticleSystem.from_numpy(host, device="cuda:0")
    a = SakuraEnsemble(ps3, Ensemble(ens1.offsets))
    b = SakuraEnsemble(ps4, Ensemble(ens1.offsets))
    a.evolve(0.25, tau)
    b.evolve_graph(0.25, tau, steps_per_graph=6)
    assert a.nsteps == b.nsteps == 16 and a.time == b.time
    for k in keys:
        assert torch.equal(getattr(ps3, k), getattr(ps4, k)), k
