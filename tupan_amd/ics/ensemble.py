"""Ensembles of independent Plummer realisations (the ensemble workload:
systems with sizes uniform in [nmin, nmax], each a seeded Plummer sphere with
its own mass function).

Each system follows ``make_plummer``'s recipe (ics/plummer.py) with its own
RandomState(seed_s): masses (then the state restore), positions, potential,
velocities, centre of mass to the origin, N-body units, virial equilibrium.
The per-system potential of all systems comes from one ensemble Phi launch
per stage instead of a kernel call per system; sums are exactly rounded
(math.fsum) per system, as the particle system's are.
"""
import math

import numpy as np
import torch

from ..lib.ensemble import Ensemble
from ..particles.system import ParticleSystem
from .imf import IMF
from .plummer import Plummer, _von_neumann


def system_sizes(nsys, nmin, nmax, seed):
    return np.random.RandomState(seed).randint(nmin, nmax + 1, size=nsys)


def system_seed(seed, s):
    return (int(seed) * 1000003 + 7919 * s) % (2**32 - 1)


def make_ensemble(nsys=4096, nmin=3, nmax=64, seed=1, imf=None, eps=0.0,
                  device=None, min_lanes=8):
    """(ParticleSystem of all systems, Ensemble plan)."""
    imf = imf if imf is not None else IMF.log_uniform(0.1, 1.0)
    sizes = system_sizes(nsys, nmin, nmax, seed)
    offsets = np.zeros(nsys + 1, dtype=np.int64)
    offsets[1:] = np.cumsum(sizes)
    ntot = int(offsets[-1])
    cols = {k: np.zeros(ntot) for k in ("mass", "eps2", "rx", "ry", "rz",
                                        "vx", "vy", "vz")}
    gens = []
    for s in range(nsys):
        n = int(sizes[s])
        g = Plummer.__new__(Plummer)
        g.n, g.eps2, g.imf, g.mfrac, g.softening_type = n, eps * eps, imf, \
            0.999, 0
        g.rng = np.random.RandomState(system_seed(seed, s))
        sl = slice(offsets[s], offsets[s + 1])
        srand = g.rng.get_state()
        mass = np.asarray(imf.sample(n, g.rng), dtype=np.float64)
        mass = mass / float(np.sum(mass))
        cols["mass"][sl] = mass
        cols["eps2"][sl] = g.set_eps2(mass)
        g.rng.set_state(srand)
        rx, ry, rz = g.set_pos(g.rng.permutation(np.arange(n)))
        cols["rx"][sl], cols["ry"][sl], cols["rz"][sl] = rx, ry, rz
        gens.append(g)
    cols["id"] = np.concatenate([np.arange(k) for k in sizes])
    ps = ParticleSystem.from_numpy(cols, device=device)
    ens = Ensemble(offsets.astype(np.int32), min_lanes=min_lanes)

    pot = ens.phi(ps).detach().cpu().numpy()
    vel = {k: np.zeros(ntot) for k in ("vx", "vy", "vz")}
    for s in range(nsys):
        sl = slice(offsets[s], offsets[s + 1])
        g = gens[s]
        rnd = _von_neumann(g.rng, g.n)
        velocity = np.sqrt(-2 * rnd * pot[sl])
        theta = np.arccos(g.rng.uniform(-1.0, 1.0, size=g.n))
        phi = g.rng.uniform(0.0, 2.0 * np.pi, size=g.n)
        vel["vx"][sl] = velocity * np.sin(theta) * np.cos(phi)
        vel["vy"][sl] = velocity * np.sin(theta) * np.sin(phi)
        vel["vz"][sl] = velocity * np.cos(theta)
    for k, v in vel.items():
        getattr(ps, k)[...] = torch.as_tensor(v, device=ps.device)

    _per_system_model(ps, ens, offsets)
    return ps, ens


def _host(ps):
    return {k: getattr(ps, k).detach().cpu().numpy().copy()
            for k in ("mass", "rx", "ry", "rz", "vx", "vy", "vz")}


def _put(ps, h):
    for k, v in h.items():
        getattr(ps, k)[...] = torch.as_tensor(v, device=ps.device)


def _pe(ps, ens, offsets):
    phi = ens.phi(ps).detach().cpu().numpy()
    m = ps.mass.detach().cpu().numpy()
    return [0.5 * math.fsum(m[a:b] * phi[a:b])
            for a, b in zip(offsets[:-1], offsets[1:])]


def _per_system_model(ps, ens, offsets):
    """com_to_origin, to_nbody_units, scale_to_virial of every system
    (body.py:157-461), with the potential from the ensemble kernel."""
    h = _host(ps)
    for a, b in zip(offsets[:-1], offsets[1:]):
        m = h["mass"][a:b]
        mt = math.fsum(m)
        for c in ("x", "y", "z"):
            cr = math.fsum(m * h["r" + c][a:b]) / mt
            cv = math.fsum(m * h["v" + c][a:b]) / mt
            h["r" + c][a:b] += -cr
            h["v" + c][a:b] += -cv
        ratio = 1.0 / mt
        h["mass"][a:b] *= ratio
        for c in ("x", "y", "z"):
            h["r" + c][a:b] *= ratio
    _put(ps, h)
    pe = _pe(ps, ens, offsets)
    h = _host(ps)
    for s, (a, b) in enumerate(zip(offsets[:-1], offsets[1:])):
        mt = math.fsum(h["mass"][a:b])
        rvir = (mt**2) / (-2 * pe[s])
        r_scale = 1.0 / rvir
        v_scale = 1 / r_scale**0.5
        for c in ("x", "y", "z"):
            h["r" + c][a:b] *= r_scale
            h["v" + c][a:b] *= v_scale
    _put(ps, h)
    pe = _pe(ps, ens, offsets)
    h = _host(ps)
    for s, (a, b) in enumerate(zip(offsets[:-1], offsets[1:])):
        m = h["mass"][a:b]
        ke = math.fsum(0.5 * m * (h["vx"][a:b]**2 + h["vy"][a:b]**2
                                  + h["vz"][a:b]**2))
        v_scale = ((-0.5 * pe[s]) / ke)**0.5
        for c in ("x", "y", "z"):
            h["v" + c][a:b] *= v_scale
    _put(ps, h)
