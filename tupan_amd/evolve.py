"""Run one integration of a Plummer sphere and report its energy error.

    python -m tupan_amd.evolve --n 16384 --method ahermite4 --eta 0.01 \\
        --t-end 1 --mode split32 [--eps 1e-4] [--imf equalmass] [--seed 1] \\
        [--graph-steps 64] [--chunk 0.0625] [--max-wall 300] \\
        [--save state.npz] [--resume state.npz] [--input ic.hdf5] \
        [--output worldline.hdf5 --dump-freq 1] [--snapshot final.hdf5]

Hermite methods run as device steps replayed in HIP graphs of
--graph-steps steps (--graph-steps 0: the host-driven driver); other
methods use the ordinary driver.  The run advances in chunks of --chunk
time units and stops at --t-end or after the first chunk that ends past
--max-wall seconds; after every chunk it prints the time reached, |dE/E| and the kinetic and
potential energies to stderr.  At the end it prints one JSON line: the time reached, steps, wall
time, |dE/E| with E from FP64 Phi and exactly rounded sums, and the force
mode.

--save writes every particle array, the time, the initial energy and the
wall time so far to an .npz file after every chunk; --resume continues a run
from such a file (the other options must be the run's own).  A step depends
only on the particle arrays and the time, and power-of-two block steps
aligned with the time never cross a chunk boundary, so a resumed run takes
the same steps as an uninterrupted one.

--input takes the initial particles from the ``Snapshot`` group of an HDF5
file in the reference's layout (tupan/simulation.py:146-147) instead of a
Plummer sphere; --output writes the worldline there, as the reference's
dumpper does (the initial state, then every dump_freq-th step of every
particle; host-driven steps only, --graph-steps 0); --snapshot writes the
final state as a snapshot.
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
    ap.add_argument("--save", default=None,
                    help="write the run's state here after every chunk")
    ap.add_argument("--resume", default=None,
                    help="continue the run saved in this file")
    ap.add_argument("--input", default=None,
                    help="initial conditions from this HDF5 snapshot")
    ap.add_argument("--output", default=None,
                    help="write the worldline to this HDF5 file")
    ap.add_argument("--dump-freq", type=int, default=1)
    ap.add_argument("--snapshot", default=None,
                    help="write the final state to this HDF5 file")
    args = ap.parse_args(argv)
    if args.output and "hermite" in args.method and args.graph_steps > 0:
        ap.error("--output needs host-driven steps (--graph-steps 0)")

    import os
    import numpy as np
    import torch
    from .diagnostics import energy
    from .ics.plummer import make_plummer
    from .integrator import Integrator
    from .lib import extensions
    from .lib.utils import ctype
    from .particles import ParticleSystem

    wall_before = 0.0
    if args.resume:
        with np.load(args.resume, allow_pickle=False) as z:
            meta = json.loads(str(z["__meta__"]))
            arrays = {k: z[k] for k in z.files if k != "__meta__"}
        ps = ParticleSystem.from_numpy(arrays, device="cuda:0")
        t = float(meta["t"])
        e0 = float(meta["e0"])
        wall_before = float(meta["wall_s"])
        print("resumed at t = %.6f (%.1f s before)" % (t, wall_before),
              file=sys.stderr, flush=True)
    else:
        if args.input:
            from .io import IO
            ps = IO(args.input, "r").load_snapshot(device="cuda:0")
            args.n = ps.n
        else:
            ps = make_plummer(args.n, args.eps, (args.imf,), seed=args.seed,
                              device="cuda:0")
        type(ps).t_curr = 0.0
        t = 0.0
        e0 = energy(ps)
        print("t = 0  K = %.15e  U = %.15e"
              % (ps.kinetic_energy, ps.potential_energy),
              file=sys.stderr, flush=True)
    ctype.force_mode = args.mode
    extensions.install("HIP", "float64")
    hermite = "hermite" in args.method
    kw = {"device_steps": True} if hermite and args.graph_steps > 0 else {}
    if args.output:
        from .io import IO
        kw.update(dumpper=IO(args.output, "a"), dump_freq=args.dump_freq)
    it = Integrator(args.eta, t, ps, method=args.method, **kw)

    def save(t_now, wall_now):
        arrays = it.particle_system.to_numpy()
        meta = {"t": t_now, "e0": e0, "wall_s": wall_now, "n": args.n,
                "eps": args.eps, "imf": args.imf, "seed": args.seed,
                "method": args.method, "eta": args.eta, "mode": args.mode}
        tmp = args.save + ".tmp.npz"
        np.savez(tmp, __meta__=np.array(json.dumps(meta)), **arrays)
        os.replace(tmp, args.save)

    torch.cuda.synchronize()
    t0 = time.perf_counter() - wall_before
    while t < args.t_end:
        target = min(args.t_end, t + args.chunk)
        if kw.get("device_steps"):
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
        if args.save:
            save(t, time.perf_counter() - t0)
        if args.max_wall and time.perf_counter() - t0 - wall_before > \
                args.max_wall:
            break
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    q = it.particle_system
    ctype.force_mode = "fp64"
    extensions.install("HIP", "float64")
    e1 = energy(q)
    steps = int(q.nstep.max().item())
    if args.snapshot:
        from .io import IO
        IO(args.snapshot, "w").dump_snapshot(q)
    print(json.dumps({"n": args.n, "eps": args.eps, "imf": args.imf,
                      "seed": args.seed, "method": args.method,
                      "eta": args.eta, "mode": args.mode, "t": t,
                      "t_end": args.t_end, "steps": steps, "wall_s": wall,
                      "steps_per_s": steps / wall if wall > 0 else None,
                      "dE_E": abs((e1 - e0) / e0)}), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
