"""Two-pass Sakura form and the learned Kepler guess on the GPU.

  * two-pass vs one-pass Sakura: the same pair drifts summed far-first,
    with the far pairs' mass ratio an ulp from IEEE division, so
    per-particle results agree to rounding amplified by the cancellation in
    the sums (<= 1e-12 of each particle's vector norm) and the work counters
    are identical; the golden
    Sakura vectors hold at the one-pass bar (1e-11) through the two-pass
    form, including lists that overflow their slots;
  * network kernels against the float64 CPU oracle (oracle/mlp.py): FP32
    forward <= 1e-5, loss and gradient <= 1e-4 of the largest gradient
    component, Adam step, BF16 inference within 2e-3 of the oracle's BF16
    emulation (MFMA accumulation order) in units of the output scale;
  * the learned guess: trained on solves logged from Sakura steps, it
    leaves the converged states unchanged (positions and velocities after
    the drift within 1e-14 per particle; the drift increments, differences
    of nearby states, within 1e-11) and its guesses are closer to the
    converged anomaly than the plain guess; the iteration counts of both
    are printed.
"""
import numpy as np
import pytest

from conftest import vec_rel_err

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

F8 = ("mass", "rx", "ry", "rz", "eps2", "vx", "vy", "vz")


def sakura_run(ps, dt, flag, path, guess=None, counters=True):
    from tupan_amd.learned import sakura_kernel
    ext = sakura_kernel(path=path, guess=guess)
    if counters:
        ext.kernel.counters = torch.zeros(5, dtype=torch.int64,
                                          device="cuda:0")
    res = ext.calc(ps, ps, dt, flag)
    out = [r.detach().cpu().numpy().copy() for r in res]
    cnt = ext.kernel.counters.cpu().tolist() if counters else None
    return out, cnt, ext.kernel.near_pairs


@pytest.fixture(scope="module")
def plummer4k():
    from oracle_backend import install_hip
    from tupan_amd.ics.plummer import make_plummer
    install_hip("float64")
    return make_plummer(4096, 0.0, ("equalmass",), seed=5, device="cuda:0")


@pytest.mark.parametrize("flag", [-1, 1, -2, 2])
def test_two_pass_equals_one_pass(plummer4k, flag):
    one, c1, n1 = sakura_run(plummer4k, 1.0 / 64, flag, "one-pass")
    two, c2, n2 = sakura_run(plummer4k, 1.0 / 64, flag, "two-pass")
    assert n1 == -1 and n2 > 0
    er, ev = vec_rel_err(two[:3], one[:3]), vec_rel_err(two[3:], one[3:])
    print("two-pass vs one-pass: dr %.2e dv %.2e" % (er, ev))
    assert er <= 1e-12 and ev <= 1e-12
    assert c1[:4] == c2[:4], (c1, c2)


def test_two_pass_golden_and_overflow(golden):
    """Golden Sakura sets through the two-pass form (24 particles at most:
    every list fits its slots; overflow is covered by the dense cluster
    below)."""
    from tupan_amd.particles import ParticleSystem
    g = golden("sakura_f64.npz")
    checked = 0
    for tag in ("nb", "r0", "r1"):
        d = {k: g["%s_%s" % (tag, k)] for k in F8}
        ps = ParticleSystem.from_numpy(d, device="cuda:0")
        for flag in (-1, 1, -2, 2):
            for dt in (1.0 / 64, 0.5):
                key = "%s_f%d_dt%g" % (tag, flag, dt)
                ref = [g["%s_%s" % (key, o)] for o in ("drx", "dry", "drz",
                                                       "dvx", "dvy", "dvz")]
                two, _, _ = sakura_run(ps, dt, flag, "two-pass")
                assert vec_rel_err(two[:3], ref[:3]) <= 1e-11, (key,)
                assert vec_rel_err(two[3:], ref[3:]) <= 1e-11, (key,)
                checked += 1
    assert checked == 24


def test_dense_cluster_overflow_equals_one_pass():
    from tupan_amd.particles import ParticleSystem
    rng = np.random.default_rng(4)
    n = 300
    d = {k: rng.normal(0, 1, n) for k in ("rx", "ry", "rz")}
    d.update({k: rng.normal(0, 0.05, n) for k in ("vx", "vy", "vz")})
    d["mass"] = np.full(n, 1.0 / n)
    d["eps2"] = np.zeros(n)
    ps = ParticleSystem.from_numpy(d, device="cuda:0")
    one, c1, _ = sakura_run(ps, 1.0 / 32, -1, "one-pass")
    two, c2, near = sakura_run(ps, 1.0 / 32, -1, "two-pass")
    assert c1[0] > 32 * n  # most pairs are Kepler pairs: lists overflow
    assert near < c1[0]
    assert vec_rel_err(two[:3], one[:3]) <= 1e-12
    assert vec_rel_err(two[3:], one[3:]) <= 1e-12
    assert c1[:4] == c2[:4]


def random_net_case(n, seed):
    rng = np.random.RandomState(seed)
    params = (rng.normal(0, 0.25, 4545)).astype(np.float32)
    feats = np.stack([np.exp(rng.uniform(-6, 1, n)), rng.normal(0, 1, n),
                      rng.uniform(0, 4, n), rng.uniform(0, 0.3, n)])
    feats = feats.astype(np.float32)
    norm = np.array([-2.5, 0.1, 1.5, 0.1, 0.6, 1.1, 0.8, 3.0, 0.02, -0.01],
                    dtype=np.float32)
    target = rng.normal(-0.01, 0.02, n).astype(np.float32)
    return params, norm, feats, target


def test_network_kernels_vs_oracle():
    from oracle import mlp
    from tupan_amd.learned import KeplerNet
    params, norm, feats, target = random_net_case(1000, 7)
    net = KeplerNet(params=params, norm=norm)
    f = torch.as_tensor(feats, device="cuda:0").contiguous()
    t = torch.as_tensor(target, device="cuda:0")
    ref = mlp.delta(params.astype(np.float64), norm, feats)
    got = net.forward(f).cpu().numpy()
    scale = np.max(np.abs(ref - norm[9]))
    assert np.max(np.abs(got - ref)) <= 1e-5 * scale
    loss, grad = net.loss_and_grad(f, t)
    rl, rg = mlp.loss_and_grad(params.astype(np.float64), norm, feats, target)
    assert abs(float(loss) - rl) <= 1e-5 * rl
    g = grad.cpu().numpy()
    assert np.max(np.abs(g - rg)) <= 1e-4 * np.max(np.abs(rg))
    # BF16 inference: FP32 sums in the kernel's order (matrix-core order in
    # the 64x64 layer) against the emulation's wide sums; an activation
    # whose FP32 value sits near a BF16 rounding boundary can round the
    # other way (one BF16 ulp, 2^-8 relative), so single items may differ
    # by up to ~1e-2 of the output scale while the typical item agrees to
    # a few 1e-4.  A wrong matrix-core operand layout gives O(1) errors.
    bf = net.infer(f).cpu().numpy()
    emu = mlp.delta_bf16(params, norm, feats)
    err = np.abs(bf - emu) / scale
    print("bf16 inference vs emulation: median %.2e max %.2e"
          % (np.median(err), err.max()))
    assert err.max() <= 1e-2 and np.median(err) <= 5e-4
    fp32_err = np.abs(bf - ref) / scale
    assert np.median(fp32_err) <= 2e-2
    # one Adam step
    p0 = params.astype(np.float64)
    net.adam_step(1e-3)
    want, _, _ = mlp.adam(p0, g.astype(np.float64), 0, 0, 1e-3, 0.9, 0.999,
                          1e-8, 1)
    np.testing.assert_allclose(net.params.cpu().numpy(), want, rtol=1e-6,
                               atol=1e-7)


def test_learned_guess_gain_and_identical_results():
    from tupan_amd.ics.plummer import make_plummer
    from tupan_amd.learned import KeplerNet, collect_trajectory
    train_ps = make_plummer(4096, 0.0, ("equalmass",), seed=11,
                            device="cuda:0")
    feats, target, iters = collect_trajectory(train_ps, 6)
    assert feats.shape[1] > 10000
    net = KeplerNet(seed=1)
    net.fit_norm(feats, target)
    hist = net.train(feats, target, epochs=6, batch=4096, lr=3e-3,
                     lr_end=3e-4)
    print("training loss per epoch", ["%.4f" % h for h in hist])
    assert hist[-1] < 0.5 * hist[0]

    test_ps = make_plummer(4096, 0.0, ("equalmass",), seed=12,
                           device="cuda:0")
    plain, cp, near = sakura_run(test_ps, 1.0 / 128, -1, "two-pass")
    learned, cl, near2 = sakura_run(test_ps, 1.0 / 128, -1, "two-pass",
                                    guess=net)
    assert near == near2 and cp[0] == cl[0]
    print("near pairs %d; Laguerre steps per solve: plain %.3f learned %.3f;"
          " fallbacks %d" % (near, cp[1] / cp[0], cl[1] / cl[0], cl[4]))
    # converged states: positions and velocities after the drift agree to
    # 1e-14; the drift increments themselves are differences p1 - p0 of
    # nearby states, so their relative agreement is looser
    r0 = [getattr(test_ps, k).cpu().numpy() for k in ("rx", "ry", "rz")]
    v0 = [getattr(test_ps, k).cpu().numpy() for k in ("vx", "vy", "vz")]
    for base, k0 in ((r0, 0), (v0, 3)):
        a = [b + d for b, d in zip(base, learned[k0:k0 + 3])]
        b = [b + d for b, d in zip(base, plain[k0:k0 + 3])]
        assert vec_rel_err(a, b) <= 1e-14
    assert vec_rel_err(learned[:3], plain[:3]) <= 1e-11
    assert vec_rel_err(learned[3:], plain[3:]) <= 1e-11
    # guess quality on the test state's own solves
    f2, t2, _ = collect_trajectory(test_ps, 1)
    pred = net.infer(f2)
    err_learned = torch.median(torch.abs(pred - t2)).item()
    err_plain = torch.median(torch.abs(t2)).item()
    print("median |delta - guess|: plain %.3e learned %.3e"
          % (err_plain, err_learned))
    assert err_learned < err_plain


@pytest.mark.parametrize("flag", [-1, 2])
def test_ensemble_learned_guess_same_states(flag):
    """The learned form of the ensemble kernel (tupan_ensemble_set_guess,
    the shipped network) on the ensemble workload (4096 systems of
    U{3..64} particles, log-uniform masses in [0.1, 1], eps = 0, seed 1),
    one pair-drift stage at dt = 1/64.
      * The same leapfrog pairs (every stage's far/near test agrees).  A
        guess can spare a pair the subdivision its plain-guess solve needs
        (at flag 2: 4 subdivisions plain, 2 learned), which changes the
        solve count; with equal subdivisions the solves are equal.
      * States after the drift within 1e-12 per particle.  Each solve stops
        after a correction of at most 2^-42 (2.3e-13) of s, so solves from
        different guesses converge that far apart, and in a few-body system
        one pair dominates a particle's drift (measured: 3e-14 in positions,
        1.3e-13 in velocities).  The increments dr = mu (r1 - r0) are
        printed, not bounded: relative to themselves they carry the states'
        difference amplified by |r| / |dr|, unbounded where a particle's
        increments nearly cancel (2.4e-7 at flag 2).
      * Fewer Laguerre steps than the plain guess (flag -1: 2.227 -> 2.009
        per solve).  Steps and fallbacks are printed."""
    import os
    from tupan_amd.ics.ensemble import make_ensemble
    from tupan_amd.learned import KeplerNet
    path = os.path.join(os.path.dirname(__file__), os.pardir, "tupan_amd",
                        "learned", "kepler_net.npz")
    net = KeplerNet.load(path, device="cuda:0")
    ps, ens = make_ensemble(nsys=4096, nmin=3, nmax=64, seed=1,
                            device="cuda:0")
    print("ensemble 4096 systems, flag %d, plain then learned" % flag,
          flush=True)
    res = {}
    for name, guess in (("plain", None), ("learned", net)):
        ens.set_guess(guess)
        ens.counters = torch.zeros(4, dtype=torch.int64, device="cuda:0")
        out = [t.cpu().numpy().copy() for t in ens.sakura(ps, 1.0 / 64, flag)]
        fb = int(ens.fallbacks.item()) if guess is not None else 0
        res[name] = (out, ens.counters.cpu().tolist(), fb)
    ens.set_guess(None)
    ens.counters = None
    (plain, cp, _), (learned, cl, fb) = res["plain"], res["learned"]
    print("solves %d; Laguerre steps per solve: plain %.4f learned %.4f; "
          "fallbacks %d" % (cp[0], cp[1] / cp[0], cl[1] / cl[0], fb))
    print("counters plain", cp, "learned", cl)
    r0 = [getattr(ps, k).cpu().numpy() for k in ("rx", "ry", "rz")]
    v0 = [getattr(ps, k).cpu().numpy() for k in ("vx", "vy", "vz")]
    errs = []
    for base, k0 in ((r0, 0), (v0, 3)):
        a = [b + d for b, d in zip(base, learned[k0:k0 + 3])]
        b = [b + d for b, d in zip(base, plain[k0:k0 + 3])]
        errs.append(vec_rel_err(a, b))
    errs += [vec_rel_err(learned[:3], plain[:3]),
             vec_rel_err(learned[3:], plain[3:])]
    print("states %.2e %.2e increments %.2e %.2e" % tuple(errs))
    for name, (g, r) in (("dr", (learned[:3], plain[:3])),
                         ("dv", (learned[3:], plain[3:]))):
        g, r = np.stack(g), np.stack(r)
        e = np.sqrt(((g - r) ** 2).sum(0)) / np.maximum(
            np.sqrt((r * r).sum(0)), 1e-300)
        top = np.argsort(e)[-5:]
        print(name, "largest", e[top].tolist(), "at", top.tolist(),
              "n > 1e-9:", int((e > 1e-9).sum()))
    assert cp[3] == cl[3]
    if cp[2] == cl[2]:
        assert cp[0] == cl[0]
    # the guess takes fewer Laguerre steps than the plain guess
    assert cl[1] < cp[1]
    assert max(errs[:2]) <= 1e-12


@pytest.mark.parametrize("flag", [-1, 2])
def test_graph_safe_two_pass_abi(plummer4k, flag):
    """tupan_sakura_graph (the two-pass form without its host read of the
    near-pair count) through the C ABI: with room for the list it equals
    the synchronising two-pass form bit for bit, work counters included;
    with a capacity below the listed count it sets status bit 2."""
    import ctypes
    from tupan_amd import _native
    from tupan_amd.lib.hip_backend import current_stream_handle
    ref, cref, near = sakura_run(plummer4k, 1.0 / 64, flag, "two-pass")
    assert near > 1
    lib = _native.lib()
    fields = _native.ptr_array([getattr(plummer4k, f).data_ptr() for f in F8])
    for cap, want in ((near, 0), (near - 1, 2)):
        outs = [torch.zeros(4096, dtype=torch.float64, device="cuda:0")
                for _ in range(6)]
        cnt = torch.zeros(5, dtype=torch.int64, device="cuda:0")
        status = torch.zeros(1, dtype=torch.int32, device="cuda:0")
        rc = lib.tupan_sakura_graph(
            4096, fields, 4096, fields, 1.0 / 64, flag,
            _native.ptr_array([o.data_ptr() for o in outs]), cap,
            cnt.data_ptr(), status.data_ptr(), current_stream_handle())
        _native.check(rc, "tupan_sakura_graph")
        torch.cuda.synchronize()
        assert int(status.item()) == want, (cap, int(status.item()))
        if want == 0:
            for o, r in zip(outs, ref):
                np.testing.assert_array_equal(o.cpu().numpy(), r)
            assert cnt.cpu().tolist()[:4] == cref[:4]
