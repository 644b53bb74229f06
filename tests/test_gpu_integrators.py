"""Drivers on the GPU against the same drivers on the CPU oracle backend,
from the same initial state (Hermite, Sakura, SIA with Kepler leaves), and
the ahermite4 run of the 1024-particle Plummer workload.

Bars: after a few steps the device and oracle trajectories agree to 1e-11
in position (the forces differ by <= 1e-13 per step; a few steps of a
collisional system amplify that modestly)."""
import numpy as np
import pytest

from oracle_backend import install_hip, install_oracle

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def cpu_plummer(n, eps, seed):
    from tupan_amd.ics.plummer import make_plummer
    install_oracle("float64")
    try:
        ps = make_plummer(n, eps, ("equalmass",), seed=seed, device="cpu")
    finally:
        install_hip("float64")
    return ps.to_numpy()


def run(method, eta, host, steps, backend, t_end=1.0, mode=None, **kw):
    from tupan_amd.integrator import Integrator
    from tupan_amd.particles import ParticleSystem
    if backend == "oracle":
        install_oracle("float64")
        dev = "cpu"
    else:
        install_hip("float64", **({"mode": mode} if mode else {}))
        dev = "cuda:0"
    try:
        ps = ParticleSystem.from_numpy(host, device=dev)
        type(ps).t_curr = 0.0
        it = Integrator(eta, 0.0, ps, method=method, **kw)
        for _ in range(steps):
            it.evolve_step(t_end)
        out = it.particle_system.to_numpy()
        t = type(ps).t_curr
    finally:
        install_hip("float64")
        type(ps).include_pn_corrections = False
    order = np.argsort(out["id"])
    return {k: v[order] for k, v in out.items()}, t


@pytest.mark.parametrize("method,eta,steps,n", [
    ("ahermite4", 0.01, 5, 256),
    ("hermite6", 1.0 / 256, 3, 128),
    ("asakura", 0.05, 3, 96),
    ("sia21h.dkd", 0.02, 2, 128),
    ("nreg", 0.02, 4, 96),
    ("anreg", 0.02, 4, 96),
])
def test_driver_gpu_matches_oracle_backend(method, eta, steps, n):
    host = cpu_plummer(n, 4.0 / n if "sakura" not in method else 0.0, 11)
    g, tg = run(method, eta, host, steps, "hip")
    c, tc = run(method, eta, host, steps, "oracle")
    if "nreg" in method:
        # NREG's physical time comes from its energy sums (1e-13 apart)
        assert abs(tg - tc) <= 1e-12 * abs(tc)
    else:
        assert tg == tc
    for k in ("rx", "ry", "rz"):
        np.testing.assert_allclose(g[k], c[k], rtol=0, atol=1e-11)
    for k in ("vx", "vy", "vz"):
        np.testing.assert_allclose(g[k], c[k], rtol=0, atol=1e-10)


def test_plummer_1024_ahermite4_energy():
    """The 1024-particle workload: Plummer, eps = 4/N, ahermite4 with
    eta = 0.01 (shared block step), to t = 1/16; |dE/E| stays small and the
    FP64 device potential matches the oracle's."""
    from oracle import Oracle
    from tupan_amd.diagnostics import energy
    from tupan_amd.ics.plummer import make_plummer
    from tupan_amd.integrator import Integrator
    install_hip("float64")
    ps = make_plummer(1024, 4.0 / 1024, ("equalmass",), seed=1,
                      device="cuda:0")
    host = ps.to_numpy()
    (phi_ref,) = Oracle("float64").phi(host, host)
    ps.set_phi(ps)
    np.testing.assert_allclose(ps.phi.cpu().numpy(), phi_ref, rtol=1e-13)
    type(ps).t_curr = 0.0
    e0 = energy(ps)
    it = Integrator(0.01, 0.0, ps, method="ahermite4")
    nstep = 0
    while it.time < 1.0 / 16:
        it.evolve_step(1.0 / 16)
        nstep += 1
    e1 = energy(it.particle_system)
    err = abs((e1 - e0) / e0)
    print("ahermite4 N=1024 to t=1/16: %d steps, |dE/E| = %.3e" % (nstep,
                                                                  err))
    assert err < 1e-12
    assert nstep > 5


@pytest.mark.parametrize("mode", ["fp64", "split32"])
@pytest.mark.parametrize("method,n", [("sia21h.dkd", 300), ("sia22h.kdk", 200),
                                      ("sia21h.kdk", 40)])
def test_sia_device_subtree_equals_host_levels(method, n, mode):
    """SIA 'h' steps with the subsets of <= 64 particles run on the device
    (one launch each) equal the host driver's level-by-level steps bit for
    bit: particle order, every array and the driver time.  n = 40 runs the
    whole step in one subtree; the larger systems mix host and device
    levels."""
    from tupan_amd.ics.imf import IMF
    from tupan_amd.ics.plummer import make_plummer
    from tupan_amd.integrator import sia
    from tupan_amd.lib import extensions
    from tupan_amd.particles import ParticleSystem
    host = make_plummer(n, 1e-3, IMF.power_law(-2.35, 0.1, 10.0), seed=9,
                        device="cuda:0").to_numpy()
    extensions.install("HIP", "float64", mode=mode)
    try:
        runs = []
        for dev in (False, True):
            ps = ParticleSystem.from_numpy(host, device="cuda:0")
            type(ps).t_curr = 0.0
            it = sia.SIA(0.05, 0.0, ps, method, device_subtree=dev)
            counts = extensions.count_pairs(True)
            for k in range(2):
                it.evolve_step(0.05 * (k + 1))
            extensions.count_pairs(False)
            runs.append((it.ps, type(ps).t_curr, dict(counts)))
    finally:
        install_hip("float64")
    (a, ta, ca), (b, tb, cb) = runs
    assert ta == tb
    # the same Tstep and Acc interactions (Kepler leaves are not counted on
    # the device)
    assert (ca.get("Tstep"), ca.get("Acc")) == (cb.get("Tstep"), cb.get("Acc"))
    for k in ("id", "mass", "rx", "ry", "rz", "vx", "vy", "vz", "time",
              "nstep", "tstep", "tstepij", "ax", "ay", "az"):
        assert torch.equal(getattr(a, k), getattr(b, k)), k
    levels = torch.unique(a.tstep).numel()
    assert levels >= 3, levels


def test_plummer_1024_ahermite4_to_t1():
    """The 1024-particle workload in full: Plummer, eps = 4/N, ahermite4
    with eta = 0.01 to t = 1; |dE/E| at t = 1 stays below 1e-11 (measured
    1.3e-13 on an MI355X) and every particle ends at t = 1."""
    from tupan_amd.diagnostics import energy
    from tupan_amd.ics.plummer import make_plummer
    from tupan_amd.integrator import Integrator
    install_hip("float64")
    ps = make_plummer(1024, 4.0 / 1024, ("equalmass",), seed=1,
                      device="cuda:0")
    type(ps).t_curr = 0.0
    e0 = energy(ps)
    it = Integrator(0.01, 0.0, ps, method="ahermite4")
    nstep = 0
    while it.time < 1.0:
        it.evolve_step(1.0)
        nstep += 1
    q = it.particle_system
    err = abs((energy(q) - e0) / e0)
    print("ahermite4 N=1024 to t=1: %d steps, |dE/E| = %.3e" % (nstep, err))
    assert it.time == 1.0
    assert torch.all(q.time == 1.0)
    assert err < 1e-11


def test_sia_pn_gpu_matches_oracle_backend():
    """SIA with PN kicks (order 7, c = 20) on a 48-body group with a compact
    binary: the HIP run (PNAcc and Acc on the device) against the same
    driver on the oracle backend."""
    rng = np.random.default_rng(4)
    n = 48
    host = {"mass": np.full(n, 1.0 / n), "eps2": np.full(n, 1e-4)}
    pos = rng.normal(0, 0.5, (n, 3))
    vel = rng.normal(0, 0.3, (n, 3))
    pos[1] = pos[0] + np.array([0.03, 0.0, 0.0])
    vel[1] = vel[0] + np.array([0.0, np.sqrt(2 / n / 0.03), 0.0])
    for k, (p, v) in enumerate(zip(("rx", "ry", "rz"), ("vx", "vy", "vz"))):
        host[p] = pos[:, k].copy()
        host[v] = vel[:, k].copy()
    for method in ("sia21s.dkd", "sia22a.kdk"):
        g, tg = run(method, 1.0 / 128, host, 4, "hip", pn_order=7,
                    clight=20.0)
        c, tc = run(method, 1.0 / 128, host, 4, "oracle", pn_order=7,
                    clight=20.0)
        assert tg == tc
        for k in ("rx", "ry", "rz"):
            np.testing.assert_allclose(g[k], c[k], rtol=0, atol=1e-11)
        for k in ("vx", "vy", "vz", "pn_ke", "pn_mvx", "pn_amz"):
            np.testing.assert_allclose(g[k], c[k], rtol=0, atol=1e-10)


@pytest.mark.parametrize("method,eta,steps", [
    ("ahermite4", 0.01, 6), ("hermite4", 1.0 / 512, 4),
    ("ahermite2", 0.005, 4), ("ahermite6", 0.01, 3), ("ahermite8", 0.01, 3),
])
def test_device_hermite_steps_match_host_driver(method, eta, steps):
    """tupan_hermite_step (step size, predictor, corrector and bookkeeping
    on the device) against the host driver on the same device forces: the
    same step sizes and time, and states within a few ulps (the host
    driver's elementwise division by a constant may be torch's reciprocal
    multiply on the GPU)."""
    host = cpu_plummer(256, 4.0 / 256, 11)
    g, tg = run(method, eta, host, steps, "hip")
    d, td = run(method, eta, host, steps, "hip", device_steps=True)
    assert tg == td
    np.testing.assert_array_equal(g["tstep"], d["tstep"])
    np.testing.assert_array_equal(g["nstep"], d["nstep"])
    np.testing.assert_array_equal(g["time"], d["time"])
    for k in ("rx", "ry", "rz", "vx", "vy", "vz"):
        np.testing.assert_allclose(d[k], g[k], rtol=1e-13, atol=1e-15)


@pytest.mark.parametrize("mode", ["fp64", "split32"])
@pytest.mark.parametrize("method,eta,steps", [
    ("ahermite4", 0.01, 6), ("hermite4", 1.0 / 512, 4),
])
def test_device_hermite4_fused_launches_match_host_driver(method, eta, steps,
                                                          mode):
    """At 257 to 1448 particles the order-4 device step runs its predictor,
    correctors and bookkeeping in the per-particle stores of its three
    AccJerk launches (predicted and once-corrected states in scratch): at
    N = 1024 against the host driver on the same device forces, the same
    step sizes and time and states within a few ulps."""
    host = cpu_plummer(1024, 4.0 / 1024, 11)
    g, tg = run(method, eta, host, steps, "hip", mode=mode)
    d, td = run(method, eta, host, steps, "hip", mode=mode, device_steps=True)
    assert tg == td
    np.testing.assert_array_equal(g["tstep"], d["tstep"])
    np.testing.assert_array_equal(g["nstep"], d["nstep"])
    np.testing.assert_array_equal(g["time"], d["time"])
    for k in ("rx", "ry", "rz", "vx", "vy", "vz", "ax", "ay", "az", "jx",
              "jy", "jz"):
        np.testing.assert_allclose(d[k], g[k], rtol=1e-13, atol=1e-15,
                                   err_msg=k)


@pytest.mark.parametrize("mode", ["fp64", "split32"])
def test_device_hermite4_start_pass_equals_separate_calls(mode):
    """The adaptive order-4 step of a small set computes Tstep and the start
    forces in one pass (OpAccJerkTstep): its a0, j0 and Tstep outputs equal
    the separate AccJerk and Tstep calls bit for bit, also where a lane's
    sums of one of the two overflow FP32 (split32: unsoftened pairs 1e-15
    apart, one approaching, one receding; only that interaction's sums are
    evaluated again in FP64, as in its own kernel).  N = 600; a step with
    t = t_end runs the pass and changes no particle."""
    from tupan_amd.integrator import Integrator
    from tupan_amd.lib import extensions
    from tupan_amd.particles import ParticleSystem
    rng = np.random.default_rng(21)
    n = 600
    d = {k: rng.random(n) for k in ("rx", "ry", "rz", "vx", "vy", "vz")}
    d["mass"] = np.full(n, 1.0 / n) * (1 + rng.random(n))
    d["eps2"] = np.zeros(n)
    x0 = (d["rx"][0], d["ry"][0], d["rz"][0])
    d["rx"][1] = x0[0] + 1e-15
    d["ry"][1], d["rz"][1] = x0[1], x0[2]
    d["rx"][2], d["ry"][2], d["rz"][2] = x0[0], x0[1] + 1e-6, x0[2]
    d["rx"][3], d["ry"][3], d["rz"][3] = x0[0] + 1e-15, x0[1] + 1e-6, x0[2]
    d["vx"][0], d["vx"][1] = 1.0, -1.0
    d["vx"][2], d["vx"][3] = -1.0, 1.0
    eta = 0.01
    print("%s start pass, N = 600 with close pairs" % mode, flush=True)
    install_hip("float64", mode=mode)
    try:
        ps = ParticleSystem.from_numpy(d, device="cuda:0")
        type(ps).t_curr = 0.0
        aj = [r.detach().cpu().numpy().copy() for r in
              extensions.AccJerk("HIP", "float64", mode=mode).calc(ps, ps)]
        ts = [r.detach().cpu().numpy().copy() for r in
              extensions.Tstep("HIP", "float64", mode=mode).calc(ps, ps, eta)]
        it = Integrator(eta, 0.0, ps, method="ahermite4", device_steps=True)
        dev = it.integrator._device()
        dev.scratch.zero_()
        dev.set_control(0.0)
        dev.launch()
        torch.cuda.synchronize()
        sc = dev.scratch.cpu().numpy()
        q = it.particle_system.to_numpy()
    finally:
        install_hip("float64")
    a0 = [sc[(6 + k) * n:(7 + k) * n] for k in range(3)]
    j0 = [sc[(9 + k) * n:(10 + k) * n] for k in range(3)]
    for k in range(3):
        np.testing.assert_array_equal(a0[k], aj[k])
        np.testing.assert_array_equal(j0[k], aj[3 + k])
    np.testing.assert_array_equal(sc[18 * n:19 * n], ts[0])
    np.testing.assert_array_equal(sc[19 * n:20 * n], ts[1])
    assert np.all(np.isfinite(np.stack(aj))) and np.all(np.isfinite(ts[0]))
    for k in ("rx", "vx", "time"):
        np.testing.assert_array_equal(q[k], d[k] if k in d else 0 * q[k])


@pytest.mark.parametrize("n", [256, 1024])
def test_device_hermite_graph_replays_equal_eager_steps(n):
    """evolve() replaying a graph of 16 device steps equals the same steps
    launched one by one, bit for bit, and stops at t_end exactly (the steps
    of the last replay past t_end change nothing); N = 1024 takes the fused
    order-4 launches."""
    from tupan_amd.integrator import Integrator
    from tupan_amd.particles import ParticleSystem
    host = cpu_plummer(n, 4.0 / n, 5)
    t_end = 0.0625
    out = []
    install_hip("float64")
    for graph in (False, True):
        ps = ParticleSystem.from_numpy(host, device="cuda:0")
        type(ps).t_curr = 0.0
        it = Integrator(0.01, 0.0, ps, method="ahermite4", device_steps=True)
        if graph:
            replays = it.integrator.evolve(t_end, steps_per_graph=16)
            assert replays >= 1
        else:
            while it.time < t_end:
                it.evolve_step(t_end)
        out.append((it.particle_system.to_numpy(), it.time))
    (a, ta), (b, tb) = out
    assert ta == tb == t_end
    for k in ("rx", "ry", "rz", "vx", "vy", "vz", "ax", "jz", "tstep",
              "tstepij", "time", "nstep"):
        np.testing.assert_array_equal(a[k], b[k], err_msg=k)


@pytest.mark.parametrize("mode", ["fp64", "split32"])
@pytest.mark.parametrize("method,eta,steps", [
    ("sia21s.dkd", 1.0 / 256, 4), ("sia21a.dkd", 0.02, 4),
    ("sia43a.kdk", 0.02, 3), ("sia45s.kdk", 1.0 / 256, 3),
    ("sia22a.dkd", 0.02, 3),
])
def test_device_sia_steps_match_host_driver(method, eta, steps, mode):
    """tupan_sia_step (the shared-step 's' and 'a' variants: step size,
    drift/kick sequence and bookkeeping on the device) against the host
    driver on the same device forces: the same step sizes and time and every
    array bit for bit (the updates are the driver's x + v * h with
    h = c_e * (c_b * tau), one rounding at a time)."""
    host = cpu_plummer(128, 4.0 / 128, 13)
    g, tg = run(method, eta, host, steps, "hip", mode=mode)
    d, td = run(method, eta, host, steps, "hip", mode=mode,
                device_steps=True)
    assert tg == td
    for k in ("rx", "ry", "rz", "vx", "vy", "vz", "ax", "ay", "az", "tstep",
              "time", "nstep"):
        np.testing.assert_array_equal(d[k], g[k], err_msg=k)
    if method[5] == "a":
        np.testing.assert_array_equal(d["tstepij"], g["tstepij"])


def test_device_sia_graph_replays_equal_eager_steps():
    """SIA.evolve replaying a graph of 8 device steps equals the same steps
    launched one by one, bit for bit, and stops at t_end exactly."""
    from tupan_amd.integrator import Integrator
    from tupan_amd.particles import ParticleSystem
    host = cpu_plummer(128, 4.0 / 128, 7)
    t_end = 0.125
    out = []
    install_hip("float64")
    for graph in (False, True):
        ps = ParticleSystem.from_numpy(host, device="cuda:0")
        type(ps).t_curr = 0.0
        it = Integrator(0.02, 0.0, ps, method="sia21a.dkd", device_steps=True)
        if graph:
            assert it.integrator.evolve(t_end, steps_per_graph=8) >= 1
        else:
            while it.time < t_end:
                it.evolve_step(t_end)
        out.append((it.particle_system.to_numpy(), it.time))
    (a, ta), (b, tb) = out
    assert ta == tb == t_end
    for k in ("rx", "ry", "rz", "vx", "vy", "vz", "ax", "tstep", "tstepij",
              "time", "nstep"):
        np.testing.assert_array_equal(a[k], b[k], err_msg=k)


def _levels_run(host, method, mode, steps, t_step, how):
    """`steps` steps of base step t_step of an 'h' method: 'host' levels,
    'device' (eagerly enqueued slots) or 'graph' (slots replayed from the
    captured graphs), one evolve_step each, or 'evolve' (SIA.evolve to each
    step's end)."""
    from tupan_amd.integrator import Integrator
    from tupan_amd.lib import extensions
    from tupan_amd.particles import ParticleSystem
    install_hip("float64", mode=mode)
    try:
        ps = ParticleSystem.from_numpy(host, device="cuda:0")
        type(ps).t_curr = 0.0
        kw = {"host": {}, "device": {"device_levels": "eager"},
              "graph": {"device_levels": True},
              "evolve": {"device_levels": True}}[how]
        it = Integrator(0.02, 0.0, ps, method=method, **kw)
        cnt = extensions.count_pairs(True)
        for k in range(steps):
            if how == "evolve":
                it.integrator.evolve(t_step * (k + 1))
            else:
                it.evolve_step(t_step * (k + 1))
        extensions.count_pairs(False)
        out = it.particle_system.to_numpy()
        t = type(ps).t_curr
    finally:
        install_hip("float64")
    return out, t, dict(cnt)


@pytest.mark.parametrize("mode", ["fp64", "split32"])
@pytest.mark.parametrize("method", ["sia21h.dkd", "sia22h.kdk"])
def test_device_levels_match_host_driver(method, mode):
    """Device-scheduled hierarchical steps (every level visit decided on the
    device, no host read of level sizes) against the host-driven levels on a
    2048-particle power-law Plummer sphere, several levels above 64
    particles: every array in the same particle order, the time and the
    Tstep/Acc interaction counts, bit for bit; graph-replayed slots equal
    eagerly enqueued ones; SIA.evolve reaches the same state as evolve_step
    when the steps end on t_end."""
    from tupan_amd.ics.imf import IMF
    from tupan_amd.ics.plummer import make_plummer
    install_hip("float64")
    host = make_plummer(2048, 1e-4, IMF.power_law(-2.35, 0.1, 10.0), seed=3,
                        device="cuda:0").to_numpy()
    g, tg, cg = _levels_run(host, method, mode, 2, 0.02 / 16, "host")
    d, td, cd = _levels_run(host, method, mode, 2, 0.02 / 16, "device")
    r, tr, cr = _levels_run(host, method, mode, 2, 0.02 / 16, "graph")
    v, tv, cv = _levels_run(host, method, mode, 2, 0.02 / 16, "evolve")
    print("%s %s: t %r, pairs host %s device %s" % (method, mode, tg, cg, cd))
    assert tg == td == tr == tv
    assert cg == cd == cr == cv
    for k in g:
        np.testing.assert_array_equal(d[k], g[k], err_msg=k)
        np.testing.assert_array_equal(r[k], g[k], err_msg=k)
        np.testing.assert_array_equal(v[k], g[k], err_msg=k)


def test_device_levels_small_set_takes_host_step():
    """A set of at most 64 particles with device_levels=True takes the host
    driver's step (one subtree launch): the same state as the host path,
    through evolve_step and through SIA.evolve."""
    from tupan_amd.ics.imf import IMF
    from tupan_amd.ics.plummer import make_plummer
    install_hip("float64")
    host = make_plummer(48, 1e-4, IMF.power_law(-2.35, 0.1, 10.0), seed=3,
                        device="cuda:0").to_numpy()
    g, tg, cg = _levels_run(host, "sia21h.dkd", "fp64", 2, 0.02 / 16, "host")
    d, td, cd = _levels_run(host, "sia21h.dkd", "fp64", 2, 0.02 / 16, "graph")
    v, tv, cv = _levels_run(host, "sia21h.dkd", "fp64", 2, 0.02 / 16, "evolve")
    assert tg == td == tv and cg == cd == cv
    for k in g:
        np.testing.assert_array_equal(d[k], g[k], err_msg=k)
        np.testing.assert_array_equal(v[k], g[k], err_msg=k)


def test_device_levels_block_step_workload():
    """The block-step workload (Plummer N = 262144, Salpeter masses,
    eps = 1e-4, seed 1, sia21h.dkd, eta = 0.02): one step of base step
    eta/64 with device-scheduled levels equals the host-driven levels bit
    for bit (every array, time, interaction counts)."""
    from tupan_amd.ics.imf import IMF
    from tupan_amd.ics.plummer import make_plummer
    install_hip("float64")
    host = make_plummer(262144, 1e-4, IMF.power_law(-2.35, 0.1, 10.0), seed=1,
                        device="cuda:0").to_numpy()
    g, tg, cg = _levels_run(host, "sia21h.dkd", "split32", 1, 0.02 / 64,
                            "host")
    d, td, cd = _levels_run(host, "sia21h.dkd", "split32", 1, 0.02 / 64,
                            "graph")
    print("N = 262144: t %r, pairs %s" % (tg, cg))
    assert tg == td and cg == cd
    for k in g:
        np.testing.assert_array_equal(d[k], g[k], err_msg=k)


def test_torch_scalar_division_on_device():
    """Records how torch divides a device tensor by a Python number (true
    division or multiplication by the reciprocal); the device Hermite step
    divides as the reference's numpy does."""
    x = torch.rand(1 << 16, dtype=torch.float64, device="cuda:0") * 10
    q = (x / 3).cpu().numpy()
    xn = x.cpu().numpy()
    true_div = np.array_equal(q, xn / 3)
    recip = np.array_equal(q, xn * (1.0 / 3))
    print("torch device x / 3: true division %s, reciprocal multiply %s"
          % (true_div, recip))
    assert true_div or recip


def test_device_hermite_bad_control_raises():
    """A non-finite end time, set from the host, makes the step-size kernel
    flag its status and the host raise; the kernels exit normally and the
    particles are untouched."""
    from tupan_amd import _native
    from tupan_amd.integrator import Integrator
    from tupan_amd.particles import ParticleSystem
    host = cpu_plummer(64, 4.0 / 64, 3)
    install_hip("float64")
    ps = ParticleSystem.from_numpy(host, device="cuda:0")
    type(ps).t_curr = 0.0
    it = Integrator(0.01, 0.0, ps, method="ahermite4", device_steps=True)
    with pytest.raises(_native.TupanHipError):
        it.evolve_step(float("nan"))
    q = it.particle_system.to_numpy()
    for k in ("rx", "vx", "time", "nstep"):
        np.testing.assert_array_equal(q[k], host[k] if k in host else q[k])
    assert int(q["nstep"].max()) == 0


# Three unsoftened bodies (found by a search with the oracle) whose Tstep at
# eta = 2 is NaN for body 1 only: its receding partner makes its w^2 sum fall
# below -1, so dt = eta / sqrt(1 + sum) has no value, while bodies 0 and 2
# get finite steps.
NAN_TSTEP_BODIES = {
    "mass": [0.0050455183982152825, 0.0029851304450925534,
             0.001907530456191219],
    "rx": [0.0, 0.05803323859868507, 0.029869613281892263],
    "ry": [0.0, 0.0, 0.006719948779563594],
    "rz": [0.0, 0.0, 0.0],
    "eps2": [0.0, 0.0, 0.0],
    "vx": [3.1669457576042443, 2.6407219741636783, 1.2667052456498304],
    "vy": [-2.2035098806466507, 0.05202897425988651, 0.6836861907765345],
    "vz": [0.0, 0.0, 0.0],
}


@pytest.mark.parametrize("mode", ["fp64", "split32"])
def test_device_hermite_nan_tstep_raises(mode):
    """A NaN Tstep of one particle stops a device ahermite step (status 2,
    TupanHipError), as the host driver's min over the steps propagates the
    NaN and fails; the particles are untouched.  A step-size reduction that
    skipped the NaN would take a step from the other two particles."""
    from oracle import Oracle
    from tupan_amd import _native
    from tupan_amd.integrator import Integrator
    from tupan_amd.particles import ParticleSystem
    host = {k: np.array(v) for k, v in NAN_TSTEP_BODIES.items()}
    ts = Oracle("float64").tstep(host, host, 2.0)[0]
    assert np.isnan(ts[1]) and np.all(np.isfinite(ts[[0, 2]]))
    install_hip("float64", mode=mode)
    try:
        ps = ParticleSystem.from_numpy(host, device="cuda:0")
        type(ps).t_curr = 0.0
        it = Integrator(2.0, 0.0, ps, method="ahermite4", device_steps=True)
        with pytest.raises(_native.TupanHipError):
            it.evolve_step(8.0)
        q = it.particle_system.to_numpy()
        for k in ("rx", "vx"):
            np.testing.assert_array_equal(q[k], host[k])
        assert int(q["nstep"].max()) == 0
        ps = ParticleSystem.from_numpy(host, device="cuda:0")
        type(ps).t_curr = 0.0
        it = Integrator(2.0, 0.0, ps, method="ahermite4")
        with pytest.raises(ValueError):
            it.evolve_step(8.0)
    finally:
        install_hip("float64")


def _sakura_run(host, t_end, mode, method="sakura", eta=1.0 / 64,
                capacity=None):
    """sakura or asakura to t_end: 'host' = the host driver, 'eager' =
    device steps one by one, 'graph' = device steps replayed as graphs of
    8.  capacity, if given, replaces the probed near-pair capacity (to force
    list overflows)."""
    from tupan_amd.integrator import Integrator
    from tupan_amd.particles import ParticleSystem
    install_hip("float64")
    ps = ParticleSystem.from_numpy(host, device="cuda:0")
    type(ps).t_curr = 0.0
    kw = {} if mode == "host" else {"device_steps": True}
    it = Integrator(eta, 0.0, ps, method=method, **kw)
    dev = None
    if mode != "host" and capacity is not None:
        dev = it.integrator._device()
        dev.capacity = capacity
    if mode == "graph":
        assert it.integrator.evolve(t_end, steps_per_graph=8) >= 1
    else:
        while it.time < t_end:
            it.evolve_step(t_end)
    out = it.particle_system.to_numpy()
    return out, it.time, (dev.capacity if dev is not None else None)


@pytest.mark.parametrize("method,eta,t_end,capacity", [
    ("sakura", 1.0 / 64, 21.0 / 64, None),
    ("sakura", 1.0 / 64, 21.0 / 64, 16),
    ("asakura", 0.05, 0.25, None),
])
def test_device_sakura_steps_match_host_driver(method, eta, t_end, capacity):
    """tupan_sakura_step (drifts, pair drifts in the graph-safe two-pass
    form, updates, the asakura step size and the bookkeeping on the device),
    one by one and replayed as HIP graphs, equals the host driver bit for bit
    on a 2048-particle Plummer sphere with eps = 0 (the host driver's calls
    take the same two-pass form at ni nj = 2^22); with a near-pair capacity
    of 16 every first try overflows, is detected from the status word,
    undone and rerun with a doubled capacity."""
    host = cpu_plummer(2048, 0.0, 7)
    ref, t_ref, _ = _sakura_run(host, t_end, "host", method, eta)
    for mode in ("eager", "graph"):
        got, t_got, cap = _sakura_run(host, t_end, mode, method, eta,
                                      capacity)
        assert t_got == t_ref
        if capacity is not None:
            assert cap > capacity
        keys = ("rx", "ry", "rz", "vx", "vy", "vz", "time", "nstep", "tstep")
        if method == "asakura":
            keys += ("tstepij",)
        for k in keys:
            np.testing.assert_array_equal(got[k], ref[k], err_msg=(mode, k))


def test_sakura_graph_replays_survive_workspace_growth():
    """A captured graph keeps the scratch workspace addresses it was
    recorded with.  Between its replays a call on a 64 times larger set
    (near-pair list beyond the recorded one) makes the library outgrow
    those workspaces; the outgrown buffers must stay valid, so the replays
    after it still equal the host driver bit for bit."""
    from tupan_amd.integrator import Integrator
    from tupan_amd.particles import ParticleSystem
    eta, t_end = 1.0 / 64, 13.0 / 64
    host = cpu_plummer(2048, 0.0, 7)
    ref, t_ref, _ = _sakura_run(host, t_end, "host", "sakura", eta)
    install_hip("float64")
    ps = ParticleSystem.from_numpy(host, device="cuda:0")
    type(ps).t_curr = 0.0
    it = Integrator(eta, 0.0, ps, method="sakura", device_steps=True)
    sak = it.integrator
    sak.initialize(t_end)
    dev = sak._device()
    dev.set_control(t_end)
    dev.probe_capacity()
    t, grown = dev.run(dev.launch)
    assert not grown
    graph = dev.capture(4)
    t, grown = dev.run(graph.replay)
    assert not grown
    # outgrow the workspaces: the two-pass probe of a 131072-particle set
    # (about 11 times the near pairs of the 2048-particle one); the device's
    # free memory shows that a workspace grew
    from tupan_amd.ics.plummer import make_plummer
    big = make_plummer(131072, 0.0, ("equalmass",), seed=3, device="cuda:0")
    it_big = Integrator(eta, 0.0, big, method="sakura", device_steps=True)
    dev_big = it_big.integrator._device()
    torch.cuda.synchronize()
    free0 = torch.cuda.mem_get_info()[0]
    dev_big.probe_capacity()
    torch.cuda.synchronize()
    assert torch.cuda.mem_get_info()[0] < free0
    del dev_big, it_big, big
    for _ in range(2):
        t, grown = dev.run(graph.replay)
        assert not grown
    torch.cuda.synchronize()
    assert t == t_ref
    got = ps.to_numpy()
    for k in ("rx", "ry", "rz", "vx", "vy", "vz", "time", "nstep", "tstep"):
        np.testing.assert_array_equal(got[k], ref[k], err_msg=k)
