"""Benchmark: all-pairs AccJerk on a Plummer sphere of N = 65536 on one GPU.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

A step is one all-pairs AccJerk evaluation (N x N interactions, self pairs
included as the reference loops over them) on device-resident inputs.  The
headline value is the split-FP32 path in Gint/s (1e9 pair interactions per
second); the FP64 path is timed the same way beside it.  Kernel time comes
from HIP events recorded by libtupan_hip.so around the all-pairs kernel on
the stream it runs on; the step time is bracketed by device synchronisation.

Also reported:
  roofline      algorithmic FLOP/s of the dominant kernel (42 flops per pair,
                the reference's count, acc_jerk_kernel_common.h:56) against
                the vector peak that bounds it
  energy        |dE/E| of a short ahermite4 run in FP64 and in split-FP32 from
                the same initial state
  sakura        Kepler solves/s of one Sakura pair-drift call (flag -1,
                dt 1/64) on a Plummer sphere of N = 16384 with eps = 0
  sakura_steps  the fixed-step sakura driver on the same sphere, 65 steps
                driven from the host and as device steps replayed in HIP
                graphs of 16, steps/s and bitwise agreement
  sakura_ensemble  4096 few-body systems, Sakura dt 1/64 to t = 1
  block_steps   sia21h.dkd hierarchical block steps on a 262144-particle
                power-law-mass Plummer sphere, pair interactions per second
                over whole steps
  other_kernels Acc, Tstep and PNAcc (order 7, c = 128) on the N = 65536
                state, split32 and fp64
  hermite       ahermite4 (eta 0.01) on a 1024-particle Plummer sphere
                (eps = 4/N) to t = 1: host-driven steps, device steps and
                device steps replayed as a HIP graph, steps/s and |dE/E|
  cpu_baseline  the CPU oracle (serial C99 restatement of the reference C
                backend, same operation order, gcc -O3 -std=c99) on a bounded
                sample of i-rows against all N j-particles, scaled per pair
"""
import argparse
import json

import numpy as np
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

FLOPS_PER_PAIR = 42          # acc_jerk_kernel_common.h:56
PEAK_FP64 = 78.6e12          # MI355X vector FP64
PEAK_FP32 = 157.3e12         # MI355X vector FP32 (counts packed FP32)
HBM_PEAK = 8.0e12
# HBM bytes per launch of the all-pairs kernel at N = 65536, from the two
# calibrated rocprofv3 counter runs in profiles/accjerk_plummer65536_pmc_
# traffic.md (FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0, KiB -> bytes).
TRAFFIC_65536 = {
    "split32": (14871.8 * 2.0 + 196608.0) * 1024,
    "fp64": (14637.1 * 2.0 + 98304.0) * 1024,
}


# Sakura far-pair kernel (flag -1, dt 1/64, Plummer N = 16384, eps = 0):
# the reference's own operation count for a far pair, one operator = one
# flop (sakura_kernel_common.h:45-123, 194-243: pair set-up and sums 36,
# pre-drift 6, far/near test 13, leapfrog 32), and the hardware counts per
# pair from rocprofv3 (SQ_INSTS_VALU x 64 / pairs, SQ_INSTS_VALU_FLOPS_FP64
# x 64 / pairs; profiles/sakura_plummer16384_far_counters.md).
SAKURA_FAR_REF_FLOPS = 87
SAKURA_FAR_HW = {"valu_per_pair": 126.5, "fp64_flops_per_pair": 127.0}
# Ensemble Sakura kernel (4096 systems, one pair-drift stage): hardware
# counts per pair drift from the same profile (nearly every pair is a
# Kepler solve, whose iterations the reference does not count).
ENSEMBLE_HW = {"valu_per_pair": 2378.8, "fp64_flops_per_pair": 2097.2}
CLOCK_PEAK = 2.4e9           # MI355X peak engine clock
SIMDS = 1024                 # 256 CUs x 4 SIMDs
FP64_CYCLES_PER_WAVE_INSTR = 4


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=65536)
    ap.add_argument("--energy-steps", type=int, default=100)
    ap.add_argument("--sakura-n", type=int, default=16384)
    ap.add_argument("--ensemble-systems", type=int, default=4096)
    ap.add_argument("--learned-net",
                    default=os.path.join(HERE, "tupan_amd", "learned",
                                         "kepler_net.npz"))
    ap.add_argument("--sia-n", type=int, default=262144,
                    help="block-step workload size (0 skips it)")
    ap.add_argument("--sia-tau", type=float, default=0.02 / 64,
                    help="base step of the timed sia21h.dkd block steps")
    ap.add_argument("--sia-steps", type=int, default=2)
    ap.add_argument("--sia-eta-step", default="",
                    help="force mode of one extra whole eta step of the "
                         "block-step workload, host and device levels "
                         "(default: none)")
    ap.add_argument("--hermite-n", type=int, default=1024,
                    help="Plummer N of the ahermite4 leg (0: skip)")
    ap.add_argument("--hermite-t-end", type=float, default=1.0,
                    help="end time of the ahermite4 leg")
    ap.add_argument("--sia-modes", default="split32,fp64")
    ap.add_argument("--ensemble-lanes", type=int, default=0,
                    help="lanes per system of the ensemble plan (0 = tuned)")
    ap.add_argument("--cpu-rows", type=int, default=0,
                    help="i-rows of the CPU baseline sample (0 = auto)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-probe-mb", type=int, default=0,
                    help="run one elementwise copy of this many MiB first "
                         "(counter calibration under rocprofv3 --pmc)")
    ap.add_argument("--modes", default="split32,fp64,ieee64",
                    help="force modes to time (the first is the headline)")
    ap.add_argument("--skip", default="",
                    help="comma-separated parts to leave out: sakura_steps, "
                         "ensemble_graph, ensemble_learned, force_error")
    return ap.parse_args()


def cpu_model():
    """The host CPU's model name (first processor in /proc/cpuinfo)."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_kernels():
    """(timed library, kind): the reference's own C backend compiled from its
    sources into oracle/_ref when that build is present ("reference"), else
    the oracle, its restatement in the same operation order ("port")."""
    from oracle import Oracle, Reference, reference_available
    if reference_available("float64"):
        return Reference("float64"), "reference"
    return Oracle("float64"), "port"


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def time_accjerk(ps, mode, steps, warmup):
    import torch
    from tupan_amd import _native
    from tupan_amd.lib import extensions
    ext = extensions.AccJerk("HIP", "float64", mode=mode)
    for _ in range(warmup):
        ext.calc(ps, ps)
    torch.cuda.synchronize()
    _native.timing(True)
    t0 = time.perf_counter()
    for _ in range(steps):
        ext.calc(ps, ps)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kms, cnt = _native.timing_read()
    _native.timing(False)
    return wall / steps, kms / 1e3 / max(cnt, 1)


def time_kernel(ps, cls_name, mode, steps, *scalars):
    """(kernel seconds per call, Gint/s) of an extension kernel on ps x ps."""
    import torch
    from tupan_amd import _native
    from tupan_amd.lib import extensions
    ext = getattr(extensions, cls_name)("HIP", "float64", mode=mode)
    ext.calc(ps, ps, *scalars)
    torch.cuda.synchronize()
    _native.timing(True)
    for _ in range(steps):
        ext.calc(ps, ps, *scalars)
    torch.cuda.synchronize()
    kms, cnt = _native.timing_read()
    _native.timing(False)
    t = kms / 1e3 / max(cnt, 1)
    return {"kernel_ms": t * 1e3, "gint_s": ps.n * ps.n / t / 1e9}


def force_error(ps, rows=32):
    """Split-FP32 AccJerk against the FP64 path on the same state, per
    particle |da_i| / |a_i| and |dj_i| / |j_i|: median, 99.9th percentile,
    maximum and root mean square; and the FP64 path against the oracle
    (the reference C backend's arithmetic) on `rows` sampled rows."""
    from oracle import Oracle
    from tupan_amd.lib import extensions
    got = {}
    for mode in ("split32", "fp64"):
        res = extensions.AccJerk("HIP", "float64", mode=mode).calc(ps, ps)
        got[mode] = [r.detach().cpu().numpy().copy() for r in res]

    def per_particle(a, b, grp):
        g = np.stack([a[k] for k in grp])
        r = np.stack([b[k] for k in grp])
        return np.sqrt(((g - r)**2).sum(0)) / np.sqrt((r * r).sum(0))

    def stats(e):
        return {"median": float(np.median(e)),
                "p99.9": float(np.percentile(e, 99.9)),
                "max": float(e.max()), "rms": float(np.sqrt(np.mean(e * e)))}

    out = {"vs": "FP64 path on the same state",
           "acc": stats(per_particle(got["split32"], got["fp64"], (0, 1, 2))),
           "jerk": stats(per_particle(got["split32"], got["fp64"],
                                      (3, 4, 5)))}
    host = ps.to_numpy()
    idx = np.random.default_rng(1).choice(ps.n, rows, replace=False)
    ref = Oracle("float64").acc_jerk({k: v[idx] for k, v in host.items()},
                                     host)
    e64 = per_particle([g[idx] for g in got["fp64"][:3]], ref[:3],
                       (0, 1, 2))
    out["fp64_vs_oracle"] = {"rows": rows, "max": float(e64.max())}
    return out


def energy_run(ps0, mode, steps):
    """ahermite4 (eta 0.01) for `steps` steps from a copy of ps0; returns
    |dE/E| with E from FP64 Phi and exactly rounded sums."""
    from tupan_amd.diagnostics import energy
    from tupan_amd.integrator import Integrator
    from tupan_amd.lib import extensions
    from tupan_amd.lib.utils import ctype
    ps = ps0.copy()
    type(ps).t_curr = 0.0
    e0 = energy(ps)
    old = ctype.force_mode
    ctype.force_mode = mode
    extensions.install("HIP", "float64")
    try:
        it = Integrator(0.01, 0.0, ps, method="ahermite4")
        for _ in range(steps):
            it.evolve_step(1.0)
        ps = it.particle_system
        t_end = type(ps).t_curr
    finally:
        ctype.force_mode = old
        extensions.install("HIP", "float64")
    e1 = energy(ps)
    return abs((e1 - e0) / e0), t_end


def sakura_rate(n):
    import torch
    from tupan_amd import _native
    from tupan_amd.ics.plummer import make_plummer
    from tupan_amd.lib import extensions
    ps = make_plummer(n, 0.0, ("equalmass",), seed=1, device="cuda:0")
    ext = extensions.Sakura("HIP", "float64")
    ext.calc(ps, ps, 1.0 / 64, -1)
    torch.cuda.synchronize()
    ext.kernel.counters = torch.zeros(4, dtype=torch.int64, device="cuda:0")
    _native.timing(True)
    t0 = time.perf_counter()
    ext.calc(ps, ps, 1.0 / 64, -1)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    each = _native.timing_read_each()
    kms = sum(each)
    _native.timing(False)
    solves, iters, retries, leap = [int(x) for x in
                                    ext.kernel.counters.cpu().tolist()]
    out = {"n": n, "pairs": n * n, "kepler_solves": solves,
           "laguerre_iterations": iters, "subdivisions": retries,
           "leapfrog_pairs": leap, "kernel_s": kms / 1e3, "call_s": wall,
           "solves_per_s": solves / (kms / 1e3),
           "iterations_per_s": iters / (kms / 1e3),
           "pairs_per_s": n * n / (kms / 1e3)}
    if len(each) == 2:
        # two-pass form: the far-pair kernel, then the rest of the call
        far_s = each[0] / 1e3
        out["far_kernel_s"] = far_s
        out["roofline"] = sakura_roofline(n * n, far_s)
    return out


def sakura_roofline(pairs, far_s):
    """Roofline of the far-pair kernel: FP64 FLOP/s by the reference's
    operation count and by the hardware's, and the VALU issue rate, against
    MI355X's vector FP64 peak (one wave instruction per 4 cycles per SIMD)."""
    alg = SAKURA_FAR_REF_FLOPS * pairs / far_s
    hw = SAKURA_FAR_HW["fp64_flops_per_pair"] * pairs / far_s
    issue = SAKURA_FAR_HW["valu_per_pair"] * pairs / 64 / far_s
    peak_issue = SIMDS * CLOCK_PEAK / FP64_CYCLES_PER_WAVE_INSTR
    return {"bound": "valu", "kernel": "sakura_far_kernel",
            "achieved": alg / 1e12, "peak": PEAK_FP64 / 1e12,
            "unit": "TFLOP/s", "frac": alg / PEAK_FP64,
            "flops_per_pair": SAKURA_FAR_REF_FLOPS,
            "flops_counted_as": "the reference's operations per far pair "
                                "(one operator, division or square root = 1)",
            "hw_fp64_flops_per_pair": SAKURA_FAR_HW["fp64_flops_per_pair"],
            "hw_achieved": hw / 1e12, "hw_frac": hw / PEAK_FP64,
            "valu_instructions_per_pair": SAKURA_FAR_HW["valu_per_pair"],
            "valu_issue_frac": issue / peak_issue,
            "traffic": None}


def sakura_steps(n, replays=4, per_graph=16):
    """The fixed-step sakura driver (eta 1/64) on a Plummer sphere of n
    (equal masses, eps = 0, seed 1), 1 + replays * per_graph steps two ways:
    driven from the host (two library calls per step, each reading its
    near-pair count back) and as device steps (one eager step, then HIP
    graphs of per_graph steps; the timed call includes the capacity probe
    and the capture).  Returns both rates and whether the final states agree
    bit for bit."""
    import torch
    from tupan_amd.ics.plummer import make_plummer
    from tupan_amd.integrator import Integrator
    from tupan_amd.particles import ParticleSystem
    host = make_plummer(n, 0.0, ("equalmass",), seed=1,
                        device="cuda:0").to_numpy()
    nsteps = 1 + replays * per_graph
    t_end = nsteps / 64.0
    out = {"n": n, "dt": 1.0 / 64, "steps": nsteps,
           "steps_per_graph": per_graph}
    states = {}
    for label in ("host", "graph"):
        ps = ParticleSystem.from_numpy(host, device="cuda:0")
        type(ps).t_curr = 0.0
        kw = {"device_steps": True} if label == "graph" else {}
        it = Integrator(1.0 / 64, 0.0, ps, method="sakura", **kw)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        if label == "graph":
            out["replays"] = it.integrator.evolve(t_end,
                                                  steps_per_graph=per_graph)
        else:
            while it.time < t_end:
                it.evolve_step(t_end)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        out[label] = {"wall_s": wall, "steps_per_s": nsteps / wall}
        states[label] = it.particle_system.to_numpy()
    out["graph_equals_host"] = all(
        (states["graph"][k] == states["host"][k]).all()
        for k in ("rx", "ry", "rz", "vx", "vy", "vz"))
    return out


def sakura_forms(n, net_path, reps=5):
    """The learned-guess workload: Plummer N = n, equal masses, eps = 0,
    seed 1; one Sakura pair-drift call (flag -1, dt = 1/64) in the one-pass
    form, the two-pass form with the plain guess, and the two-pass form
    with the trained network (if `net_path` exists).  Best of `reps` calls
    per form, the forms' calls interleaved so that clock changes over the
    run reach every form alike; kernel time from HIP events around the
    whole call (both passes, the near-pair compaction and the host read of
    the near-pair count)."""
    import torch
    from tupan_amd import _native
    from tupan_amd.ics.plummer import make_plummer
    from tupan_amd.learned import KeplerNet, SolveLogBuffer, sakura_kernel
    ps = make_plummer(n, 0.0, ("equalmass",), seed=1, device="cuda:0")
    forms = [("one_pass", "one-pass", None), ("two_pass", "two-pass", None)]
    if net_path and os.path.exists(net_path):
        forms.append(("two_pass_learned", "two-pass",
                      KeplerNet.load(net_path)))
    out = {"n": n, "dt": 1.0 / 64, "flag": -1}
    results = {}
    exts = {}
    for name, path, net in forms:
        exts[name] = sakura_kernel(path=path, guess=net)
        exts[name].calc(ps, ps, 1.0 / 64, -1)
    bests = {}
    for rep in range(reps):
        for name, path, net in forms:
            ext = exts[name]
            ext.kernel.counters = torch.zeros(5, dtype=torch.int64,
                                              device="cuda:0")
            torch.cuda.synchronize()
            _native.timing(True)
            t0 = time.perf_counter()
            res = ext.calc(ps, ps, 1.0 / 64, -1)
            torch.cuda.synchronize()
            wall = time.perf_counter() - t0
            kms, _ = _native.timing_read()
            _native.timing(False)
            if name not in bests or kms < bests[name][0]:
                bests[name] = (kms, wall)
            if rep == reps - 1:
                results[name] = [r.detach().cpu().numpy().copy()
                                 for r in res]
    for name, path, net in forms:
        ext = exts[name]
        best = bests[name]
        c = ext.kernel.counters.cpu().tolist()
        out[name] = {"kernel_s": best[0] / 1e3, "call_s": best[1],
                     "kepler_solves": c[0], "laguerre_iterations": c[1],
                     "iterations_per_solve": c[1] / max(1, c[0]),
                     "guess_fallbacks": c[4], "near_pairs":
                         ext.kernel.near_pairs,
                     "solves_per_s": c[0] / (best[0] / 1e3)}
        if path == "two-pass":
            # one more call with a solve log: the Laguerre steps of every
            # near pair, and the mean over waves of 64 pairs of the slowest
            # pair's steps (a wave runs as long as its slowest lane)
            log = SolveLogBuffer(ext.kernel.near_pairs + 64)
            ext.kernel.solve_log = log
            ext.calc(ps, ps, 1.0 / 64, -1)
            ext.kernel.solve_log = None
            k = ext.kernel.near_pairs
            steps = log.iterations[:k].cpu().numpy()
            waves = steps[:(k // 64) * 64].reshape(-1, 64)
            out[name]["wave_max_iterations"] = float(waves.max(1).mean())
            out[name]["max_iterations"] = int(steps.max())
    ref = results["one_pass"]

    def rel(a):
        num = sum(np.max(np.sqrt(sum((a[g + k] - ref[g + k])**2
                                     for k in range(3))) /
                         np.maximum(np.sqrt(sum(ref[g + k]**2
                                                for k in range(3))), 1e-300))
                  for g in (0, 3))
        return float(num)

    for name in results:
        if name != "one_pass":
            out[name]["max_rel_diff_vs_one_pass"] = rel(results[name])
    out["two_pass_speedup"] = out["one_pass"]["kernel_s"] / \
        out["two_pass"]["kernel_s"]
    if "two_pass_learned" in out:
        out["learned_speedup_vs_two_pass"] = out["two_pass"]["kernel_s"] / \
            out["two_pass_learned"]["kernel_s"]
        out["learned_target"] = {
            "speedup": 1.5,
            "stopping_rule": "a solve ends after a Laguerre step whose "
                             "correction is at most 2^-42 of s, so it takes "
                             "at least 2 steps unless the guess is already "
                             "that close",
            "near_pass_share": "Kepler solves are the near pass only; the "
                               "far-pair leapfrog kernel is the rest of the "
                               "call"}
    return out


def ensemble_rate(nsys, lanes, graph=True, net_path=None):
    """The ensemble workload: nsys Plummer systems with sizes U[3,64] and
    log-uniform masses in [0.1, 1], Sakura with dt = 1/64 to t = 1, every
    pair-drift stage one launch over all systems.  Timed twice from the same
    initial state: eager launches (kernel time from HIP events, counters)
    and replays of one captured 8-step HIP graph (wall time; the final
    states are compared).  With `net_path`, the eager run again from the
    same state with the learned Kepler guess (kernel time, Laguerre steps,
    fallbacks, and the final states' largest difference from the plain
    run's, per particle relative to its vector)."""
    import torch
    from tupan_amd import _native
    from tupan_amd.ics.ensemble import make_ensemble
    from tupan_amd.integrator.ensemble import SakuraEnsemble
    from tupan_amd.tune import tuned_ensemble_lanes
    if not lanes:
        lanes = tuned_ensemble_lanes()
    t0 = time.perf_counter()
    ps, ens = make_ensemble(nsys=nsys, nmin=3, nmax=64, seed=1,
                            device="cuda:0", min_lanes=lanes)
    t_ic = time.perf_counter() - t0
    ps_init = ps.copy()
    ps_learned = ps.copy() if net_path and os.path.exists(net_path) else None
    drv = SakuraEnsemble(ps, ens)
    e0 = drv.energies()
    # eager launches: kernel time (HIP events) and work counters
    ens.counters = torch.zeros(4, dtype=torch.int64, device="cuda:0")
    torch.cuda.synchronize()
    _native.timing(True)
    t0 = time.perf_counter()
    drv.evolve(1.0, 1.0 / 64)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kms, launches = _native.timing_read()
    _native.timing(False)
    counts = ens.counters.cpu().tolist()
    ens.counters = None
    # the same 64 steps as replays of one captured 8-step HIP graph
    wall_graph = same = None
    if graph:
        gdrv = SakuraEnsemble(ps_init, ens)
        g = gdrv.capture(1.0 / 64, 8)
        ens.reset_status()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(drv.nsteps // 8):
            g.replay()
        torch.cuda.synchronize()
        wall_graph = time.perf_counter() - t0
        ens.check_status()
        same = all(torch.equal(getattr(ps, k), getattr(ps_init, k))
                   for k in ("rx", "ry", "rz", "vx", "vy", "vz"))
    learned = None
    if ps_learned is not None:
        from tupan_amd.learned import KeplerNet
        ens.set_guess(KeplerNet.load(net_path, device="cuda:0"))
        ldrv = SakuraEnsemble(ps_learned, ens)
        ens.counters = torch.zeros(4, dtype=torch.int64, device="cuda:0")
        torch.cuda.synchronize()
        _native.timing(True)
        ldrv.evolve(1.0, 1.0 / 64)
        torch.cuda.synchronize()
        lkms, _ = _native.timing_read()
        _native.timing(False)
        lc = [int(x) for x in ens.counters.cpu().tolist()]
        ens.counters = None
        fb = int(ens.fallbacks.item())
        ens.set_guess(None)
        diff = 0.0
        for group in (("rx", "ry", "rz"), ("vx", "vy", "vz")):
            a = torch.stack([getattr(ps_learned, k) for k in group])
            b = torch.stack([getattr(ps, k) for k in group])
            rel = (a - b).norm(dim=0) / b.norm(dim=0).clamp_min(1e-300)
            diff = max(diff, float(rel.max()))
        learned = {"kernel_s": lkms / 1e3,
                   "kepler_solves": lc[0], "laguerre_iterations": lc[1],
                   "subdivisions": lc[2], "leapfrog_pairs": lc[3],
                   "iterations_per_solve": lc[1] / max(lc[0], 1),
                   "guess_fallbacks": fb,
                   "speedup_vs_plain": kms / lkms,
                   "final_state_max_rel_diff": diff,
                   "network": os.path.basename(net_path)}
    solves, iters, retries, leap = [int(x) for x in counts]
    sizes = np.diff(ens.offsets)
    pairs = int(np.sum(sizes.astype(np.int64)**2))
    per_launch = kms / 1e3 / max(launches, 1)
    hw = ENSEMBLE_HW["fp64_flops_per_pair"] * pairs / per_launch
    issue = ENSEMBLE_HW["valu_per_pair"] * pairs / 64 / per_launch
    roofline = {"bound": "valu", "kernel": "ensemble_sakura_kernel",
                "achieved": hw / 1e12, "peak": PEAK_FP64 / 1e12,
                "unit": "TFLOP/s", "frac": hw / PEAK_FP64,
                "flops_counted_as": "hardware FP64 FLOPs per pair drift "
                                    "(rocprofv3 SQ_INSTS_VALU_FLOPS_FP64)",
                "flops_per_pair": ENSEMBLE_HW["fp64_flops_per_pair"],
                "valu_instructions_per_pair": ENSEMBLE_HW["valu_per_pair"],
                "valu_issue_frac": issue / (SIMDS * CLOCK_PEAK /
                                            FP64_CYCLES_PER_WAVE_INSTR),
                "launch_s": per_launch, "traffic": None}
    return {"systems": nsys, "particles": int(ens.ntot),
            "sum_n2": pairs, "steps": drv.nsteps, "dt": 1.0 / 64,
            "lanes_per_system_min": lanes, "launches": launches,
            "kepler_solves": solves, "laguerre_iterations": iters,
            "subdivisions": retries, "leapfrog_pairs": leap,
            "kernel_s": kms / 1e3, "wall_s": wall,
            "wall_s_graph": wall_graph, "graph_equals_eager": same,
            "ic_s": t_ic,
            "solves_per_s": solves / (kms / 1e3),
            "pair_drifts_per_s": 2 * drv.nsteps * pairs / (kms / 1e3),
            "iterations_per_solve": iters / max(solves, 1),
            "iterations_per_s": iters / (kms / 1e3),
            "roofline": roofline,
            "learned_guess": learned,
            "dE_E_quantiles": energy_quantiles(drv, e0)}


def energy_quantiles(drv, e0):
    """|dE/E| over the systems: median, 90th, 99th percentile and maximum.
    Sakura's pairwise splitting is not exact for three-body encounters, so
    a few few-body systems with eps = 0 lose energy conservation (the
    reference algorithm on the CPU oracle gives median 4e-5 and maximum 0.6
    over the first 600 of these systems)."""
    import torch
    e1 = drv.energies()
    err = torch.abs((e1 - e0) / e0).cpu().numpy()
    return {"median": float(np.median(err)),
            "p90": float(np.percentile(err, 90)),
            "p99": float(np.percentile(err, 99)),
            "max": float(err.max())}


def block_step_rate(n, tau, steps, modes, device_levels=True):
    """The block-step workload: Plummer sphere of n particles with masses
    dN/dm ~ m^-2.35 on [0.1, 10] (normalised to total mass 1), eps = 1e-4,
    seed 1, integrated by sia21h.dkd (eta = 0.02, hierarchical power-of-two
    slow/fast levels) with base step `tau`: one untimed step, then `steps`
    timed ones.  Throughput counts every n_i * n_j of every kernel call of
    the timed steps (Tstep on each level, Acc for kicks and slow<->fast
    coupling) over the wall time."""
    import torch
    from tupan_amd import _native
    from tupan_amd.diagnostics import energy
    from tupan_amd.ics.imf import IMF
    from tupan_amd.ics.plummer import make_plummer
    from tupan_amd.integrator import Integrator
    from tupan_amd.lib import extensions
    from tupan_amd.lib.utils import ctype
    ps0 = make_plummer(n, 1e-4, IMF.power_law(-2.35, 0.1, 10.0), seed=1,
                       device="cuda:0")
    out = {"n": n, "imf": "dN/dm ~ m^-2.35 on [0.1, 10]", "eps": 1e-4,
           "seed": 1, "method": "sia21h.dkd", "eta": 0.02, "tau": tau,
           "steps": steps}
    old = ctype.force_mode
    for mode in modes:
        log("bench: block steps, %s" % mode)
        ps = ps0.copy()
        type(ps).t_curr = 0.0
        e0 = energy(ps)
        ctype.force_mode = mode
        extensions.install("HIP", "float64")
        try:
            it = Integrator(0.02, 0.0, ps, method="sia21h.dkd")
            it.evolve_step(tau)
            torch.cuda.synchronize()
            cnt = extensions.count_pairs(True)
            _native.timing(True)
            t0 = time.perf_counter()
            for k in range(steps):
                it.evolve_step(tau * (k + 2))
            torch.cuda.synchronize()
            wall = time.perf_counter() - t0
            kms, launches = _native.timing_read()
            _native.timing(False)
            extensions.count_pairs(False)
            ps = it.particle_system
        finally:
            ctype.force_mode = old
            extensions.install("HIP", "float64")
        pairs = int(sum(cnt.values()))
        out[mode] = {"gint_s": pairs / wall / 1e9, "pairs": pairs,
                     "pairs_by_kernel": {k: int(v) for k, v in cnt.items()},
                     "wall_s": wall, "pairwise_kernel_s": kms / 1e3,
                     "pairwise_launches": launches,
                     "dE_E": abs((energy(ps) - e0) / e0)}
        if device_levels:
            log("bench: block steps, %s, device-scheduled levels" % mode)
            out[mode]["device_levels"] = block_steps_device_levels(
                ps0, mode, tau, steps, ps.to_numpy())
    return out


def block_steps_device_levels(ps0, mode, tau, steps, host_final):
    """The same steps with every level visit scheduled on the device
    (SIA(..., device_levels=True)): one untimed step (it captures the two
    slot graphs), then `steps` timed ones, each replaying the graphs until
    the device reports the step done, no host read of level sizes.  Returns the
    rate, the level visits above 64 particles and working slots per step,
    the graph replays, and whether the final state equals the host-driven
    levels' bit for bit."""
    import torch
    from tupan_amd.diagnostics import energy
    from tupan_amd.integrator import Integrator
    from tupan_amd.lib import extensions
    from tupan_amd.lib.utils import ctype
    ps = ps0.copy()
    type(ps).t_curr = 0.0
    old = ctype.force_mode
    ctype.force_mode = mode
    extensions.install("HIP", "float64")
    try:
        it = Integrator(0.02, 0.0, ps, method="sia21h.dkd", device_levels=True)
        it.evolve_step(tau)
        torch.cuda.synchronize()
        lv = it.integrator._device_levels()
        cnt = extensions.count_pairs(True)
        replays = 0
        visits = ops = 0
        t0 = time.perf_counter()
        for k in range(steps):
            replays += lv.step(tau * (k + 2))
            visits += lv.last_visits
            ops += lv.last_ops
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        extensions.count_pairs(False)
        energy(it.particle_system)        # phi, as the host leg's dE/E left it
        final = it.particle_system.to_numpy()
    finally:
        ctype.force_mode = old
        extensions.install("HIP", "float64")
    pairs = int(sum(cnt.values()))
    same = all(np.array_equal(final[k], host_final[k]) for k in host_final)
    return {"gint_s": pairs / wall / 1e9, "pairs": pairs, "wall_s": wall,
            "level_visits_above_64_per_step": visits / steps,
            "slots_per_step": ops / steps, "graph_replays": replays,
            "slots_per_replay": lv.slots,
            "equals_host_levels": same}


def block_eta_step(n, mode):
    """One whole base step (eta = 0.02, every level of the hierarchy) of the
    block-step workload from t = 0, with host-driven levels and with
    device-scheduled levels (graphs captured inside the timed step), forces
    and Tstep in `mode`: interactions, rate, level visits, slots and graph
    replays, and whether the two final states are equal bit for bit.  About
    two minutes per driver in split32; not part of the default run."""
    import torch
    from tupan_amd.ics.imf import IMF
    from tupan_amd.ics.plummer import make_plummer
    from tupan_amd.integrator import Integrator
    from tupan_amd.lib import extensions
    from tupan_amd.lib.utils import ctype
    ps0 = make_plummer(n, 1e-4, IMF.power_law(-2.35, 0.1, 10.0), seed=1,
                       device="cuda:0")
    out = {"n": n, "method": "sia21h.dkd", "eta": 0.02, "tau": 0.02,
           "mode": mode}
    finals = {}
    old = ctype.force_mode
    for how in ("host_levels", "device_levels"):
        log("bench: one eta step, %s, %s" % (mode, how))
        ps = ps0.copy()
        type(ps).t_curr = 0.0
        ctype.force_mode = mode
        extensions.install("HIP", "float64")
        try:
            kw = {"device_levels": True} if how == "device_levels" else {}
            it = Integrator(0.02, 0.0, ps, method="sia21h.dkd", **kw)
            torch.cuda.synchronize()
            cnt = extensions.count_pairs(True)
            t0 = time.perf_counter()
            if how == "device_levels":
                it.integrator.is_initialized = True
                lv = it.integrator._device_levels()
                replays = lv.step(0.02)
            else:
                it.evolve_step(0.02)
            torch.cuda.synchronize()
            wall = time.perf_counter() - t0
            extensions.count_pairs(False)
            finals[how] = it.particle_system.to_numpy()
        finally:
            ctype.force_mode = old
            extensions.install("HIP", "float64")
        pairs = int(sum(cnt.values()))
        res = {"gint_s": pairs / wall / 1e9, "pairs": pairs, "wall_s": wall,
               "t": type(ps).t_curr}
        if how == "device_levels":
            res.update({"level_visits_above_64": lv.last_visits,
                        "slots": lv.last_ops, "graph_replays": replays,
                        "slots_per_replay": lv.slots})
        out[how] = res
    out["device_equals_host"] = all(
        np.array_equal(finals["host_levels"][k], finals["device_levels"][k])
        for k in finals["host_levels"])
    return out


def _best(fn, reps):
    best = None
    for _ in range(reps):
        t0 = time.perf_counter()
        fn()
        dt = time.perf_counter() - t0
        best = dt if best is None else min(best, dt)
    return best


def cpu_sakura(n, rows):
    """Serial Sakura (flag -1, dt = 1/64) of the reference C backend (or
    the oracle) on `rows` i-rows of the Plummer N = n state against all n j;
    Kepler solves/s and pairs/s.  The solve count is the oracle's on the
    same sample (it makes the reference's branch decisions)."""
    from oracle import Oracle
    from tupan_amd.ics.plummer import make_plummer
    host = make_plummer(n, 0.0, ("equalmass",), seed=1,
                        device="cuda:0").to_numpy()
    sub = {k: v[:rows] for k, v in host.items()}
    lib, kind = cpu_kernels()
    t = _best(lambda: lib.sakura(sub, host, 1.0 / 64, -1), 2)
    orc = Oracle("float64")
    orc.counters_reset()
    orc.sakura(sub, host, 1.0 / 64, -1)
    solves = orc.counters()[0]
    return {"value": solves / t / 1e6, "unit": "M Kepler solves/s",
            "pairs_per_s": rows * n / t, "cores": 1, "kind": kind,
            "cpu_model": cpu_model(),
            "sample": "%d i-rows x %d j, Sakura flag -1, dt 1/64, best of 2 "
                      "(%.2f s, %d solves)" % (rows, n, t, solves)}


def cpu_block_steps(n, rows, mix):
    """Serial Tstep and Acc of the reference C backend (or the oracle) on
    `rows` i-rows of the block-step workload's initial state against all n
    j; the step's interaction mix
    `mix` ({"Tstep": pairs, "Acc": pairs}) at those rates, in Gint/s."""
    from oracle import Oracle
    from tupan_amd.ics.imf import IMF
    from tupan_amd.ics.plummer import make_plummer
    host = make_plummer(n, 1e-4, IMF.power_law(-2.35, 0.1, 10.0), seed=1,
                        device="cuda:0").to_numpy()
    sub = {k: v[:rows] for k, v in host.items()}
    lib, kind = cpu_kernels()
    tt = _best(lambda: lib.tstep(sub, host, 0.02), 1)
    ta = _best(lambda: lib.acc(sub, host), 1)
    rt, ra = rows * n / tt, rows * n / ta
    pt, pa = mix.get("Tstep", 0), mix.get("Acc", 0)
    step_s = pt / rt + pa / ra
    return {"value": (pt + pa) / step_s / 1e9, "unit": "Gint/s", "cores": 1,
            "kind": kind, "cpu_model": cpu_model(),
            "tstep_gint_s": rt / 1e9, "acc_gint_s": ra / 1e9,
            "sample": "%d i-rows x %d j of Tstep (%.2f s) and Acc (%.2f s), "
                      "weighted by the step's interaction counts"
                      % (rows, n, tt, ta)}


def cpu_ensemble(nsys):
    """Serial Sakura pair drifts of the reference C backend (or the oracle)
    (flags -1 and +1, dt 1/128) on the first `nsys` systems of the ensemble
    workload's initial state, system by system; Kepler solves/s (solve
    count from the oracle on the same calls)."""
    from oracle import Oracle
    from tupan_amd.ics.ensemble import make_ensemble
    ps, ens = make_ensemble(nsys=4096, nmin=3, nmax=64, seed=1,
                            device="cuda:0")
    host = ps.to_numpy()
    off = ens.offsets
    subs = [{k: v[off[s]:off[s + 1]] for k, v in host.items()}
            for s in range(nsys)]
    lib, kind = cpu_kernels()

    def run(impl):
        for sub in subs:
            for flag in (-1, 1):
                impl.sakura(sub, sub, 1.0 / 128, flag)
    t = _best(lambda: run(lib), 2)
    orc = Oracle("float64")
    orc.counters_reset()
    run(orc)
    solves = orc.counters()[0]
    return {"value": solves / t / 1e6, "unit": "M Kepler solves/s",
            "cores": 1, "kind": kind, "cpu_model": cpu_model(),
            "sample": "first %d of 4096 systems, 2 pair-drift calls each, "
                      "best of 2 (%.2f s, %d solves)" % (nsys, t, solves)}


def hermite_run(n, eta, t_end, host_steps):
    """Plummer N = n (equal masses, eps = 4/n, seed 1), ahermite4 with
    shared block steps to t_end, three ways: steps driven by the host
    (torch elementwise updates, one library call per force; timed over the
    first host_steps steps), device steps launched one call per step, and
    the same device steps replayed as a HIP graph of 64 steps.  Returns the
    rates, the step count and |dE/E| at t_end."""
    import torch
    from tupan_amd.ics.plummer import make_plummer
    from tupan_amd.integrator import Integrator
    from tupan_amd.particles import ParticleSystem
    host = make_plummer(n, 4.0 / n, ("equalmass",), seed=1,
                        device="cuda:0").to_numpy()
    out = {"n": n, "eps": 4.0 / n, "method": "ahermite4", "eta": eta,
           "t_end": t_end, "seed": 1}
    states = {}
    for label in ("graph", "device", "host"):
        ps = ParticleSystem.from_numpy(host, device="cuda:0")
        type(ps).t_curr = 0.0
        e0 = ps.kinetic_energy + ps.potential_energy
        kw = {} if label == "host" else {"device_steps": True}
        it = Integrator(eta, 0.0, ps, method="ahermite4", **kw)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        if label == "graph":
            it.integrator.evolve(t_end, steps_per_graph=64)
        else:
            k = 0
            while it.time < t_end and (label == "device" or k < host_steps):
                it.evolve_step(t_end)
                k += 1
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        q = it.particle_system
        steps = int(q.nstep[0].item())
        out[label] = {"steps": steps, "wall_s": wall,
                      "steps_per_s": steps / wall}
        if label != "host":
            e1 = q.kinetic_energy + q.potential_energy
            out[label]["dE_E"] = abs((e1 - e0) / e0)
            states[label] = q.to_numpy()
    out["steps"] = out["graph"]["steps"]
    out["graph_equals_device"] = all(
        (states["graph"][k] == states["device"][k]).all()
        for k in ("rx", "ry", "rz", "vx", "vy", "vz"))
    return out


def cpu_hermite(n, eta):
    """Serial force work of one ahermite4 step (Tstep + 3 AccJerk on all
    n x n pairs) of the Hermite workload's initial state, by the reference
    C backend (or the oracle); steps/s."""
    from tupan_amd.ics.plummer import make_plummer
    host = make_plummer(n, 4.0 / n, ("equalmass",), seed=1,
                        device="cuda:0").to_numpy()
    lib, kind = cpu_kernels()

    def step():
        lib.tstep(host, host, eta)
        for _ in range(3):
            lib.acc_jerk(host, host)
    t = _best(step, 3)
    return {"value": 1.0 / t, "unit": "steps/s", "cores": 1, "kind": kind,
            "cpu_model": cpu_model(),
            "sample": "Tstep + 3 AccJerk of the N = %d state, best of 3 "
                      "(%.3f s per step)" % (n, t)}


def cpu_baseline(ps, rows):
    """`rows` i-rows against all j on one core, pairs/s -> Gint/s: the
    reference's own C backend (acc_jerk_kernel.c compiled from its sources
    with its cffi flags into oracle/_ref) when that build is present, else
    the oracle (its restatement, same operation order and flags); and the
    oracle's OpenMP build on all host cores."""
    import numpy as np
    from oracle import Oracle, Reference, reference_available
    host = ps.to_numpy()
    n = len(host["mass"])
    sub = {k: v[:rows] for k, v in host.items()}
    kind = "reference" if reference_available("float64") else "port"
    orc = Reference("float64") if kind == "reference" else Oracle("float64")
    best = None
    for _ in range(3):
        t0 = time.perf_counter()
        orc.acc_jerk(sub, host)
        dt = time.perf_counter() - t0
        best = dt if best is None else min(best, dt)
    rate = rows * n / best
    omp = Oracle("float64", omp=True)
    threads = omp.threads()
    rows_omp = min(n, rows * threads)
    subo = {k: v[:rows_omp] for k, v in host.items()}
    best_o = None
    for _ in range(2):
        t0 = time.perf_counter()
        omp.acc_jerk(subo, host)
        dt = time.perf_counter() - t0
        best_o = dt if best_o is None else min(best_o, dt)
    rate_o = rows_omp * n / best_o
    del np
    return ({"value": rate / 1e9, "unit": "Gint/s", "cores": 1,
             "kind": kind, "cpu_model": cpu_model(),
             "sample": "%d i-rows x %d j of the same Plummer state, FP64, "
                       "best of 3 (%.2f s)" % (rows, n, best)},
            {"value": rate_o / 1e9, "unit": "Gint/s", "cores": threads,
             "kind": "port", "cpu_model": cpu_model(),
             "sample": "%d i-rows x %d j, OpenMP over i, best of 2 (%.2f s)"
                       % (rows_omp, n, best_o)})


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(--gpus %d requested)" % args.gpus, file=sys.stderr)
        return 2
    import torch
    from tupan_amd import _native
    from tupan_amd.ics.plummer import make_plummer
    assert torch.cuda.is_available(), "no HIP device"
    cu, wg = _native.device_info()
    n = args.n
    if args.copy_probe_mb > 0:
        nel = args.copy_probe_mb * (1 << 20) // 8
        a = torch.ones(nel, dtype=torch.float64, device="cuda:0")
        b = torch.empty_like(a)
        torch.mul(a, 1.0, out=b)
        torch.cuda.synchronize()
        log("bench: copy probe %d MiB read + %d MiB written"
            % (args.copy_probe_mb, args.copy_probe_mb))
        del a, b
    log("bench: building Plummer N=%d eps=1e-4 seed=1" % n)
    ps = make_plummer(n, 1e-4, ("equalmass",), seed=1, device="cuda:0")
    pairs = float(n) * n

    modes = args.modes.split(",")
    skip = set(x for x in args.skip.split(",") if x)
    res = {}
    for mode in modes:
        k = args.steps if mode != "ieee64" else max(2, args.steps // 4)
        log("bench: %s AccJerk, %d steps" % (mode, k))
        res[mode] = time_accjerk(ps, mode, k, args.warmup)
    nan = (float("nan"), float("nan"))
    t_split, k_split = res.get("split32", nan)
    t_fp64, k_fp64 = res.get("fp64", nan)
    t_ieee, k_ieee = res.get("ieee64", nan)

    out = {
        "metric": "Gint/s AccJrk N=64k split-FP32 vs FP64 at |dE/E|<=1e-10; "
                  "Sakura solves/s",
        "value": pairs / t_split / 1e9,
        "unit": "Gint/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": t_split * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32-split (f64 sums)",
        "data": "synthetic",
        "config": {"workload": "plummer_accjerk", "n": n, "eps": 1e-4,
                   "imf": "equalmass", "seed": 1,
                   "step": "one all-pairs AccJerk evaluation"},
        "fp64": {"gint_s": pairs / t_fp64 / 1e9, "ms_per_step": t_fp64 * 1e3,
                 "kernel_ms": k_fp64 * 1e3},
        "ieee64": {"gint_s": pairs / t_ieee / 1e9,
                   "ms_per_step": t_ieee * 1e3, "kernel_ms": k_ieee * 1e3},
        "split32": {"gint_s": pairs / t_split / 1e9,
                    "ms_per_step": t_split * 1e3, "kernel_ms": k_split * 1e3},
        "split_over_fp64": t_fp64 / t_split,
        "device": {"cu": cu, "target_workgroups": wg},
    }
    flops = FLOPS_PER_PAIR * pairs
    out["roofline"] = {
        "bound": "valu",
        "bound_note": "vector-ALU compute: the pair interaction is not a "
                      "contraction, so no MFMA; peak is the published FP32 "
                      "vector peak, achieved counts the reference's 42 flops "
                      "per pair",
        "kernel": "pair_kernel<OpSplitAccJerk>",
        "achieved": flops / k_split / 1e12,
        "peak": PEAK_FP32 / 1e12,
        "unit": "TFLOP/s",
        "frac": flops / k_split / PEAK_FP32,
        "traffic": TRAFFIC_65536["split32"] if n == 65536 else None,
        "traffic_unit": "bytes per launch (rocprofv3 FETCH_SIZE, WRITE_SIZE)",
        "algorithmic_bytes": (8 + 6) * 8 * n,
        "fp64_kernel": {"kernel": "pair_kernel<OpAccJerk<Fast64>>",
                        "achieved": flops / k_fp64 / 1e12,
                        "peak": PEAK_FP64 / 1e12,
                        "frac": flops / k_fp64 / PEAK_FP64},
    }

    if "force_error" not in skip and "split32" in res and "fp64" in res:
        log("bench: split-FP32 force error against FP64")
        out["force_error"] = force_error(ps)

    out["other_kernels"] = {
        "%s_%s" % (name.lower(), mode): time_kernel(ps, name, mode, 5, *sc)
        for name, sc in (("Acc", ()), ("Tstep", (0.01,)))
        for mode in ("split32", "fp64")}
    from tupan_amd.lib import extensions as _ext
    _ext.clight.pn_order, _ext.clight.clight = 7, 128.0
    for mode in ("split32", "fp64"):
        out["other_kernels"]["pnacc_order7_%s" % mode] = time_kernel(
            ps, "PNAcc", mode, 3)

    if args.energy_steps > 0:
        log("bench: energy check, %d ahermite4 steps per mode"
            % args.energy_steps)
        e64, t64 = energy_run(ps, "fp64", args.energy_steps)
        e32, t32 = energy_run(ps, "split32", args.energy_steps)
        out["energy"] = {"method": "ahermite4", "eta": 0.01,
                         "steps": args.energy_steps, "t_end": t64,
                         "dE_E_fp64": e64, "dE_E_split32": e32}

    if args.sakura_n > 0:
        log("bench: Sakura N=%d" % args.sakura_n)
        out["sakura"] = sakura_rate(args.sakura_n)

    if args.sakura_n > 0:
        log("bench: Sakura forms and learned guess N=%d" % args.sakura_n)
        out["sakura_forms"] = sakura_forms(args.sakura_n, args.learned_net)
        if "sakura_steps" not in skip:
            log("bench: Sakura driver steps N=%d, host-driven and graphs"
                % args.sakura_n)
            out["sakura_steps"] = sakura_steps(args.sakura_n)

    if args.ensemble_systems > 0:
        log("bench: Sakura ensemble of %d systems" % args.ensemble_systems)
        out["sakura_ensemble"] = ensemble_rate(
            args.ensemble_systems, args.ensemble_lanes,
            graph="ensemble_graph" not in skip,
            net_path=None if "ensemble_learned" in skip else args.learned_net)

    if args.sia_n > 0:
        log("bench: block steps N=%d" % args.sia_n)
        out["block_steps"] = block_step_rate(
            args.sia_n, args.sia_tau, args.sia_steps,
            [m for m in args.sia_modes.split(",") if m])

    if args.sia_eta_step:
        out["block_eta_step"] = block_eta_step(args.sia_n or 262144,
                                               args.sia_eta_step)

    if args.hermite_n > 0:
        log("bench: ahermite4 N=%d to t=%g, host, device and graph steps"
            % (args.hermite_n, args.hermite_t_end))
        out["hermite"] = hermite_run(args.hermite_n, 0.01, args.hermite_t_end,
                                     2000)

    if not args.no_cpu:
        rows = args.cpu_rows or 2048
        log("bench: CPU baseline, %d rows" % rows)
        one, allc = cpu_baseline(ps, rows)
        out["cpu_baseline"] = one
        out["cpu_baseline_all_cores"] = allc
        if args.sakura_n > 0:
            log("bench: CPU baseline, Sakura")
            out["sakura"]["cpu_baseline"] = cpu_sakura(args.sakura_n, 256)
        if args.ensemble_systems > 0:
            log("bench: CPU baseline, ensemble")
            out["sakura_ensemble"]["cpu_baseline"] = cpu_ensemble(256)
        if args.hermite_n > 0:
            log("bench: CPU baseline, Hermite")
            out["hermite"]["cpu_baseline"] = cpu_hermite(args.hermite_n, 0.01)
        if args.sia_n > 0:
            log("bench: CPU baseline, block steps")
            mode = [m for m in args.sia_modes.split(",") if m][-1]
            out["block_steps"]["cpu_baseline"] = cpu_block_steps(
                args.sia_n, 64, out["block_steps"][mode]["pairs_by_kernel"])
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
