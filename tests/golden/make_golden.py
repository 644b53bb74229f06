"""Write the golden fixtures of tests/golden from the reference C backend.

Runs only where the reference checkout is present: `make -C oracle ref`
compiles tupan/lib/src/{phi,acc,acc_jerk,snap_crackle,tstep,sakura,
kepler_solver,pnacc}_kernel.c and nreg_kernels.c (FP64 and FP32 builds, the
reference's cffi flags)
into oracle/_ref, and this script calls those libraries through ctypes on
seeded inputs.  The fixtures are inputs and outputs only.

    python tests/golden/make_golden.py

Fixtures (numpy .npz, float64 unless noted):
  pairwise_f64.npz / pairwise_f32.npz
      two random particle sets A (97 particles) and B (61 particles), masses
      U[0,1), positions and velocities U[0,10), accelerations and jerks
      U[-1,1) for SnapCrackle, eps2 either 0 or U[0,0.01); B shares one
      particle position with A (a coincident pair) and A contains a
      duplicated particle; outputs of Phi, Acc, AccJerk, SnapCrackle,
      Tstep(eta=1/64), NREG X and NREG V (dt=1/64) and PNAcc (order 7,
      c = 128, the reference's own test values) for (A,B), (B,A) and (A,A),
      and for A against every prefix B[:k], k in PREFIXES; PNAcc of every
      order 0..8 for (A,B) and (A,A) at c = 128 (keys pnacc_o<order>) and of
      order 7 at c = 16 (keys pnacc_c16).
  sakura_f64.npz
      Sakura dr, dv for flags -2, -1, 0, 1, 2 and dt in {1/64, 0.5} on a
      random 24-particle set (eps2 = 0 and eps2 > 0) and on a bound 8-body
      system.
  kepler_f64.npz
      kepler_solver_kernel on two-body states: bound, unbound, near
      parabolic, long-step and softened cases.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import Reference, SPECS  # noqa: E402

PREFIXES = (1, 2, 3, 7, 16, 33, 61)
SCALAR_KERNELS = ("tstep", "nreg_x", "nreg_v")   # eta or dt = 1/64
PN_ORDER, PN_CLIGHT = 7, 128.0   # tupan/tests/test_extensions.py:110-111


def pn_scalars(order, c):
    """(order, 1/c, ..., 1/c^7) as the reference's Clight forms them
    (extensions.py:51-60)."""
    inv1 = 1.0 / c
    return (order,) + tuple(inv1**k for k in range(1, 8))
FIELDS = ("mass", "rx", "ry", "rz", "eps2", "vx", "vy", "vz",
          "ax", "ay", "az", "jx", "jy", "jz")


def particle_set(rng, n, soft):
    d = {"mass": rng.random(n)}
    for k in ("rx", "ry", "rz", "vx", "vy", "vz"):
        d[k] = rng.random(n) * 10
    for k in ("ax", "ay", "az", "jx", "jy", "jz"):
        d[k] = rng.random(n) * 2 - 1
    d["eps2"] = rng.random(n) * 0.01 if soft else np.zeros(n)
    return d


def pairwise_fixture(prec):
    dtype = np.float64 if prec == "float64" else np.float32
    ref = Reference(prec)
    rng = np.random.default_rng(20240611)
    out = {}
    for soft in (0, 1):
        a = particle_set(rng, 97, soft)
        b = particle_set(rng, 61, soft)
        # a duplicated particle in A and a particle of B on top of one of A
        for k in FIELDS:
            a[k][10] = a[k][3]
        for k in ("rx", "ry", "rz"):
            b[k][5] = a[k][20]
        a = {k: v.astype(dtype) for k, v in a.items()}
        b = {k: v.astype(dtype) for k, v in b.items()}
        tag = "s%d" % soft
        for k in FIELDS:
            out["%s_A_%s" % (tag, k)] = a[k]
            out["%s_B_%s" % (tag, k)] = b[k]
        for name in SPECS:
            if name == "sakura":
                continue
            if name == "pnacc":
                sc = pn_scalars(PN_ORDER, PN_CLIGHT)
            else:
                sc = (1.0 / 64,) if name in SCALAR_KERNELS else ()
            for lab, ii, jj in (("AB", a, b), ("BA", b, a), ("AA", a, a)):
                res = ref.pairwise(name, ii, jj, *sc)
                for o, r in zip(SPECS[name][3], res):
                    out["%s_%s_%s_%s" % (tag, name, lab, o)] = r
            for k in PREFIXES:
                bk = {f: v[:k] for f, v in b.items()}
                res = ref.pairwise(name, a, bk, *sc)
                for o, r in zip(SPECS[name][3], res):
                    out["%s_%s_pre%d_%s" % (tag, name, k, o)] = r
        for lab, ii, jj in (("AB", a, b), ("AA", a, a)):
            for order in range(9):
                res = ref.pairwise("pnacc", ii, jj,
                                   *pn_scalars(order, PN_CLIGHT))
                for o, r in zip(SPECS["pnacc"][3], res):
                    out["%s_pnacc_o%d_%s_%s" % (tag, order, lab, o)] = r
            res = ref.pairwise("pnacc", ii, jj, *pn_scalars(7, 16.0))
            for o, r in zip(SPECS["pnacc"][3], res):
                out["%s_pnacc_c16_%s_%s" % (tag, lab, o)] = r
    return out


def sakura_fixture():
    ref = Reference("float64")
    rng = np.random.default_rng(99)
    out = {}
    sets = {}
    for soft in (0, 1):
        d = particle_set(rng, 24, soft)
        d["mass"] = d["mass"] * 0.1
        sets["r%d" % soft] = d
    # a bound 8-body system: small, cold, concentrated
    n = 8
    d = {"mass": np.full(n, 1.0 / n), "eps2": np.zeros(n)}
    for k in ("rx", "ry", "rz"):
        d[k] = rng.normal(0, 0.3, n)
    for k in ("vx", "vy", "vz"):
        d[k] = rng.normal(0, 0.4, n)
    sets["nb"] = d
    for lab, d in sets.items():
        for k in FIELDS[:8]:
            out["%s_%s" % (lab, k)] = d[k]
        for flag in (-2, -1, 0, 1, 2):
            for dt in (1.0 / 64, 0.5):
                res = ref.sakura(d, d, dt, flag)
                for o, r in zip(SPECS["sakura"][3], res):
                    out["%s_f%d_dt%g_%s" % (lab, flag, dt, o)] = r
    return out


def kepler_cases():
    """(m0, m1, e2 each, r (3), v (3) relative, dt) two-body states."""
    cases = []
    # circular, elliptic, hyperbolic, near parabolic, long step, softened
    cases.append((0.5, 0.5, 0.0, [1, 0, 0], [0, 1, 0], 0.3))
    cases.append((0.7, 0.3, 0.0, [1, 0.2, 0], [0.1, 0.6, 0.2], 2.5))
    cases.append((0.5, 0.5, 0.0, [1, 0, 0], [0, 2.0, 0.3], 1.0))
    # near parabolic (alpha ~ 2e-3; exactly parabolic makes the reference's
    # (cosh t - 1)/z cancel to noise, universal_kepler_solver.h:58-61)
    cases.append((0.5, 0.5, 0.0, [1, 0, 0], [0, 1.415, 0], 4.0))
    cases.append((0.9, 0.1, 0.0, [0.5, -0.5, 0.1], [0.3, 0.9, -0.1], 37.0))
    cases.append((0.5, 0.5, 0.01, [0.2, 0.1, 0], [0.1, 1.5, 0], 0.7))
    cases.append((0.5, 0.5, 0.02, [1, 0, 0], [0, 0.9, 0.1], -1.3))
    cases.append((0.6, 0.4, 0.0, [2, 1, -1], [-0.4, 0.1, 0.2], -0.8))
    return cases


def kepler_fixture():
    ref = Reference("float64")
    out = {}
    for c, (m0, m1, e2, r, v, dt) in enumerate(kepler_cases()):
        ps = {"mass": np.array([m0, m1]), "eps2": np.array([e2, e2]),
              "rx": np.array([r[0] * 0.5, -r[0] * 0.5]) + 0.1,
              "ry": np.array([r[1] * 0.5, -r[1] * 0.5]) - 0.2,
              "rz": np.array([r[2] * 0.5, -r[2] * 0.5]),
              "vx": np.array([v[0] * 0.5, -v[0] * 0.5]) + 0.05,
              "vy": np.array([v[1] * 0.5, -v[1] * 0.5]),
              "vz": np.array([v[2] * 0.5, -v[2] * 0.5]) - 0.01}
        res = ref.kepler(ps, dt)
        for k, val in ps.items():
            out["c%d_%s" % (c, k)] = val
        out["c%d_dt" % c] = np.array(dt)
        for o, r1 in zip(("rx", "ry", "rz", "vx", "vy", "vz"), res):
            out["c%d_out_%s" % (c, o)] = r1
    out["ncases"] = np.array(len(kepler_cases()))
    return out


def main():
    np.savez_compressed(os.path.join(HERE, "pairwise_f64.npz"),
                        **pairwise_fixture("float64"))
    np.savez_compressed(os.path.join(HERE, "pairwise_f32.npz"),
                        **pairwise_fixture("float32"))
    np.savez_compressed(os.path.join(HERE, "sakura_f64.npz"),
                        **sakura_fixture())
    np.savez_compressed(os.path.join(HERE, "kepler_f64.npz"),
                        **kepler_fixture())
    for f in sorted(os.listdir(HERE)):
        if f.endswith(".npz"):
            print(f, os.path.getsize(os.path.join(HERE, f)))


if __name__ == "__main__":
    main()
