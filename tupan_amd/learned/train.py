"""Train the Kepler-guess network on solves logged from Sakura runs.

    python -m tupan_amd.learned.train [--out PATH] [--n 16384] [--seeds 2,3,4,5]
                                      [--steps 10] [--epochs 20]

Data: for every seed, a Plummer sphere of n particles (equal masses, eps = 0,
and a second one with the power-law mass function dN/dm ~ m^-2.35 on
[0.1, 10] and eps = 1e-4) is advanced by `steps` Sakura steps (dt = 1/64)
with a solve log on every pair-drift call; each near pair's first solve
gives one (features, delta) sample.  The network is trained with the
library's kernels (tupan_mlp_grad, tupan_mlp_adam) and saved as an .npz of
its FP32 parameters and normalisation.  A held-out state (seed 1, the bench
workload) reports the guess error and the Laguerre steps per solve with and
without the network.
"""
import argparse
import os
import sys
import time

import torch

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_OUT = os.path.join(HERE, "kepler_net.npz")


def log(msg):
    print(msg, flush=True)


def build_dataset(n, seeds, steps, dt):
    from ..ics.imf import IMF
    from ..ics.plummer import make_plummer
    from . import collect_trajectory
    feats, targs, iters = [], [], []
    for seed in seeds:
        for imf, eps in ((("equalmass",), 0.0),
                         (IMF.power_law(-2.35, 0.1, 10.0), 1e-4)):
            ps = make_plummer(n, eps, imf, seed=seed, device="cuda:0")
            f, t, it = collect_trajectory(ps, steps, dt)
            feats.append(f)
            targs.append(t)
            iters.append(it)
            log("  seed %d eps %g: %d solves, %.3f Laguerre steps each"
                % (seed, eps, t.numel(), float(it.double().mean())))
    return torch.cat(feats, 1), torch.cat(targs), torch.cat(iters)


def evaluate(net, n, dt):
    """Guess error and Laguerre steps per solve on the held-out state."""
    from ..ics.plummer import make_plummer
    from . import collect_trajectory, sakura_kernel
    ps = make_plummer(n, 0.0, ("equalmass",), seed=1, device="cuda:0")
    out = {}
    for name, guess in (("plain", None), ("learned", net)):
        ext = sakura_kernel(path="two-pass", guess=guess)
        ext.kernel.counters = torch.zeros(5, dtype=torch.int64,
                                          device="cuda:0")
        ext.calc(ps, ps, dt, -1)
        c = ext.kernel.counters.cpu().tolist()
        out[name] = {"solves": c[0], "steps_per_solve": c[1] / max(1, c[0]),
                     "fallbacks": c[4]}
    f, t, _ = collect_trajectory(ps, 1, dt)
    pred = net.infer(f)
    out["median_abs_delta"] = float(torch.median(torch.abs(t)))
    out["median_abs_guess_error"] = float(torch.median(torch.abs(pred - t)))
    return out


def main(argv=None):
    from . import KeplerNet
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=DEFAULT_OUT)
    ap.add_argument("--n", type=int, default=16384)
    ap.add_argument("--seeds", default="2,3,4,5")
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--dt", type=float, default=1.0 / 64)
    ap.add_argument("--epochs", type=int, default=20)
    ap.add_argument("--batch", type=int, default=16384)
    ap.add_argument("--lr", type=float, default=3e-3)
    ap.add_argument("--lr-end", type=float, default=1e-4)
    args = ap.parse_args(argv)
    seeds = [int(s) for s in args.seeds.split(",") if s]
    t0 = time.time()
    log("collecting solves: n=%d seeds=%s steps=%d" % (args.n, seeds,
                                                       args.steps))
    feats, target, iters = build_dataset(args.n, seeds, args.steps, args.dt)
    log("%d samples in %.1f s" % (target.numel(), time.time() - t0))
    net = KeplerNet(seed=0)
    net.fit_norm(feats, target)
    t1 = time.time()
    net.train(feats, target, epochs=args.epochs, batch=args.batch,
              lr=args.lr, lr_end=args.lr_end,
              callback=lambda ep, l: log("  epoch %d loss %.5f" % (ep, l)))
    torch.cuda.synchronize()
    log("trained in %.1f s" % (time.time() - t1))
    net.save(args.out)
    log("saved %s" % args.out)
    log("held-out: %s" % evaluate(net, args.n, args.dt))
    return 0


if __name__ == "__main__":
    sys.exit(main())
