"""Autotuner: ``python -m tupan_amd.tune [--quick]`` on the target GPU.

For every pairwise kernel and arithmetic mode it times each compiled tile
shape at 2, 4, 8 and 16 workgroups per CU on a uniform random cluster of one
representative size per size bucket (1024, 16384 and 65536 particles, i =
j), keeps the fastest (GPU time of the whole call, best of 3), and for the
ensemble plan it times each lanes-per-system packing on the 4096-system
workload.  The choices go to tupan_amd/tune/cache.json, which the library
applies when it loads.
"""
import argparse
import json
import sys
import time

import numpy as np
import torch

from .. import _native
from . import CACHE, load, save

KERNELS = {  # name -> (kernel id, modes)
    "phi": (_native.K_PHI, ("fp64",)),
    "acc": (_native.K_ACC, ("fp64", "split32")),
    "acc_jerk": (_native.K_ACC_JERK, ("fp64", "split32", "ieee64")),
    "snap_crackle": (_native.K_SNAP_CRACKLE, ("fp64",)),
    "tstep": (_native.K_TSTEP, ("fp64", "split32")),
    "pnacc": (_native.K_PNACC, ("fp64", "split32")),
}
SLOT = {"fp64": 0, "split32": 1, "ieee64": 2}
BUCKET_N = (1024, 16384, 65536)


def cluster(n, seed=0):
    from ..particles import ParticleSystem
    rng = np.random.default_rng(seed)
    d = {k: rng.normal(0, 1, n) for k in ("rx", "ry", "rz", "vx", "vy",
                                         "vz", "ax", "ay", "az", "jx", "jy",
                                         "jz")}
    d["mass"] = np.full(n, 1.0 / n)
    d["eps2"] = np.full(n, 1e-8)
    return ParticleSystem.from_numpy(d, device="cuda:0")


def time_call(fn, reps):
    """Best of `reps` GPU times of fn(), from events on the launch stream
    around the whole call (all-pairs kernel, split combine and the gaps
    between their launches)."""
    fn()
    torch.cuda.synchronize()
    best = None
    for _ in range(reps):
        t0 = torch.cuda.Event(enable_timing=True)
        t1 = torch.cuda.Event(enable_timing=True)
        t0.record()
        fn()
        t1.record()
        t1.synchronize()
        t = t0.elapsed_time(t1)
        best = t if best is None else min(best, t)
    return best


def tune_pairwise(quick, names):
    from ..lib import extensions
    lib = _native.lib()
    out = []
    sizes = BUCKET_N[:2] if quick else BUCKET_N
    # PNAcc at the reference's own test values (order 7, c = 128)
    extensions.clight.pn_order, extensions.clight.clight = 7, 128.0
    for b, n in enumerate(sizes):
        ps = cluster(n)
        for name, (kid, modes) in KERNELS.items():
            if name not in names:
                continue
            cls = {"phi": extensions.Phi, "acc": extensions.Acc,
                   "acc_jerk": extensions.AccJerk,
                   "snap_crackle": extensions.SnapCrackle,
                   "tstep": extensions.Tstep,
                   "pnacc": extensions.PNAcc}[name]
            for mode in modes:
                ext = cls("HIP", "float64", mode=mode)
                args = (0.01,) if name == "tstep" else ()
                results = []
                for shape in range(6):
                    for wg in (2, 4, 8, 16):
                        _native.check(lib.tupan_tune_set(kid, SLOT[mode], b,
                                                         shape, wg),
                                      "tupan_tune_set")
                        t = time_call(lambda: ext.calc(ps, ps, *args), 3)
                        results.append((t, shape, wg))
                results.sort()
                t, shape, wg = results[0]
                default = [r for r in results if r[1] == 0 and r[2] == 4][0]
                _native.check(lib.tupan_tune_set(kid, SLOT[mode], b, shape,
                                                 wg), "tupan_tune_set")
                out.append({"kernel": name, "mode": mode, "bucket": b,
                            "n": n, "shape": shape, "wg_per_cu": wg,
                            "ms": t, "default_ms": default[0]})
                print("%-12s %-7s n=%-6d shape %d wg/cu %d  %.3f ms "
                      "(default %.3f)" % (name, mode, n, shape, wg, t,
                                          default[0]), flush=True)
    return out


def tune_ensemble(nsys):
    from ..ics.ensemble import make_ensemble
    from ..lib.ensemble import Ensemble
    ps, ens = make_ensemble(nsys=nsys, nmin=3, nmax=64, seed=1,
                            device="cuda:0")
    res = []
    for lanes in (8, 16, 32, 64):
        e = Ensemble(ens.offsets, min_lanes=lanes)
        t = time_call(lambda: e.sakura(ps, 1.0 / 64, -1), 3)
        res.append((t, lanes))
        print("ensemble min_lanes %2d  %.3f ms" % (lanes, t), flush=True)
    res.sort()
    return {"min_lanes": res[0][1], "ms": res[0][0],
            "all": {str(L): t for t, L in res}}


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--quick", action="store_true")
    ap.add_argument("--ensemble-systems", type=int, default=4096)
    ap.add_argument("--kernels", default=",".join(KERNELS),
                    help="comma-separated subset; the cache keeps the other "
                         "kernels' entries")
    args = ap.parse_args(argv)
    assert torch.cuda.is_available(), "the autotuner runs on the GPU"
    names = [k for k in args.kernels.split(",") if k]
    for k in names:
        if k not in KERNELS:
            raise SystemExit("unknown kernel %r" % k)
    cu, _ = _native.device_info()
    t0 = time.time()
    old = load()
    keep = [e for e in old.get("pairwise", []) if e.get("kernel") not in names]
    data = {"device": {"cu": cu, "name": torch.cuda.get_device_name(0)},
            "pairwise": keep + tune_pairwise(args.quick, names)}
    if args.ensemble_systems > 0:
        data["ensemble"] = tune_ensemble(args.ensemble_systems)
    elif "ensemble" in old:
        data["ensemble"] = old["ensemble"]
    data["seconds"] = time.time() - t0
    save(data)
    print("wrote", CACHE)
    return 0


if __name__ == "__main__":
    sys.exit(main())
