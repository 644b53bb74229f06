"""Run one integration of a Plummer sphere and report its energy error.

    python -m tupan_amd.evolve --n 16384 --method ahermite4 --eta 0.01 \\
        --t-end 1 --mode split32 [--eps 1e-4] [--imf equalmass] [--seed 1] \\
        [--graph-steps 64] [--chunk 0.0625] [--max-wall 300]

Hermite methods run as device steps replayed in HIP graphs of
--graph-steps steps (--graph-steps 0: the host-driven driver); other
methods use the ordinary driver.  The run advances in chunks of --chunk
time units and stops at --t-end or after the first chunk that ends past
--max-wall seconds; after every chunk it prints the time reached, |dE/E| and the kinetic and
potential energies to stderr.  At the end it prints one JSON line: the time reached, steps, wall
time, |dE/E| with E from FP64 Phi and exactly rounded sums, and the force
mode.
"""
import argparse
import json
import sys
import time


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=16384)
    ap.add_argument("--eps", type=float, default=1e-4)
    ap.add_argument("--imf", default="equalmass")
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--method", default="ahermite4")
    ap.add_argument("--eta", type=float, default=0.01)
    ap.add_argument("--t-end", type=float, default=1.0)
    ap.add_argument("--mode", default="fp64",
                    choices=("fp64", "split32", "ieee64"))
    ap.add_argument("--graph-steps", type=int, default=64)
    ap.add_argument("--chunk", type=float, default=1.0 / 16)
    ap.add_argument("--max-wall", type=float, default=0.0,
                    help="stop after the chunk that passes this many "
                         "seconds (0: no limit)")
    args = ap.parse_args(argv)

    import torch
    from .diagnostics import energy
    from .ics.plummer import make_plummer
    from .integrator import Integrator
    from .lib import extensions
    from .lib.utils import ctype

    ps = make_plummer(args.n, args.eps, (args.imf,), seed=args.seed,
                      device="cuda:0")
    type(ps).t_curr = 0.0
    e0 = energy(ps)
    print("t = 0  K = %.15e  U = %.15e" % (ps.kinetic_energy, ps.potential_energy),
          file=sys.stderr, flush=True)
    ctype.force_mode = args.mode
    extensions.install("HIP", "float64")
    hermite = "hermite" in args.method
    kw = {"device_steps": True} if hermite and args.graph_steps > 0 else {}
    it = Integrator(args.eta, 0.0, ps, method=args.method, **kw)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    t = 0.0
    while t < args.t_end:
        target = min(args.t_end, t + args.chunk)
        if kw:
            it.integrator.evolve(target, steps_per_graph=args.graph_steps)
        else:
            while it.time < target:
                it.evolve_step(target)
        t = it.time
        # Phi is FP64 in every force mode, so this is the FP64 energy
        q = it.particle_system
        ke, pe = q.kinetic_energy, q.potential_energy
        print("t = %.6f  %.1f s  |dE/E| = %.3e  K = %.15e  U = %.15e"
              % (t, time.perf_counter() - t0, abs((ke + pe - e0) / e0), ke, pe),
              file=sys.stderr, flush=True)
        if args.max_wall and time.perf_counter() - t0 > args.max_wall:
            break
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    q = it.particle_system
    ctype.force_mode = "fp64"
    extensions.install("HIP", "float64")
    e1 = energy(q)
    steps = int(q.nstep.max().item())
    print(json.dumps({"n": args.n, "eps": args.eps, "imf": args.imf,
                      "seed": args.seed, "method": args.method,
                      "eta": args.eta, "mode": args.mode, "t": t,
                      "t_end": args.t_end, "steps": steps, "wall_s": wall,
                      "steps_per_s": steps / wall if wall > 0 else None,
                      "dE_E": abs((e1 - e0) / e0)}), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
