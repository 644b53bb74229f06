"""Plummer sphere initial conditions (tupan/ics/plummer.py:21-221).

The random-number call sequence is the reference's (numpy's legacy MT19937
stream, seeded once):

  1. save the state; draw the masses from the IMF; normalise them to total
     mass 1; set eps2 = eps^2/2 per particle (softening_type 0); restore the
     state, so the mass draws do not move the positions (plummer.py:95-112);
  2. positions: permutation(n), random(n), uniform(-1,1,n), uniform(0,2pi,n)
     (plummer.py:65-75);
  3. phi = Phi(ps, ps) on the device (plummer.py:121);
  4. velocities: von Neumann rejection of random() pairs (r1, r2), keeping
     r2 when r2 < r1, until n are kept, then uniform(-1,1,n),
     uniform(0,2pi,n) (plummer.py:77-93).  The scalar loop is evaluated in
     batches and the generator then advanced by exactly the number of draws
     the loop consumes, which leaves the same stream state;
  5. centre of mass to the origin, N-body units (total mass 1, virial radius
     1; eps2 is not rescaled), virial equilibrium 2K + U = 0
     (plummer.py:129-133, body.py:404-461).

The potential comes from the installed extension kernels (the HIP backend in
production), so realisations are reproducible but not bitwise equal to the
reference's host computation.
"""
import logging

import numpy as np
import torch

from ..particles.system import ParticleSystem
from .imf import IMF

logger = logging.getLogger(__name__)


def _von_neumann(rng, n):
    """n accepted r2 values of the reference's rejection loop, advancing rng
    by exactly the draws that loop makes."""
    state = rng.get_state()
    out = []
    need = n
    consumed = 0
    while need > 0:
        k = 2 * need + 64
        u = rng.random_sample(2 * k).reshape(k, 2)
        idx = np.nonzero(u[:, 1] < u[:, 0])[0]
        if len(idx) >= need:
            take = idx[:need]
            consumed += 2 * (int(take[-1]) + 1)
            out.append(u[take, 1])
            need = 0
        else:
            consumed += 2 * k
            out.append(u[idx, 1])
            need -= len(idx)
    rng.set_state(state)
    rng.random_sample(consumed)
    return np.concatenate(out) if out else np.empty(0)


class Plummer(object):
    def __init__(self, n, eps, imf, seed=None, mfrac=0.999, softening_type=0,
                 device=None):
        self.n = n
        self.eps2 = eps * eps
        self.imf = imf
        self.mfrac = mfrac
        self.softening_type = softening_type
        self.ps = ParticleSystem(n, device=device)
        self.rng = np.random.RandomState(seed)

    def set_eps2(self, mass):
        n = self.n
        if self.softening_type == 0:
            eps2 = np.ones(n)
        elif self.softening_type == 1:
            eps2 = mass**2
        elif self.softening_type == 2:
            eps2 = mass / n
        elif self.softening_type == 3:
            eps2 = (mass / n**2)**(2.0 / 3)
        elif self.softening_type == 4:
            eps2 = (1.0 / (mass * n**2))**2
        else:
            raise ValueError("Unexpected value for softening_type: {}."
                             .format(self.softening_type))
        eps2 *= self.eps2 / np.mean(eps2)
        return eps2 / 2

    def set_pos(self, irand):
        n = self.n
        rng = self.rng
        mrand = (irand + rng.random_sample(n)) * self.mfrac / n
        radius = 1.0 / np.sqrt(np.power(mrand, -2.0 / 3.0) - 1.0)
        theta = np.arccos(rng.uniform(-1.0, 1.0, size=n))
        phi = rng.uniform(0.0, 2.0 * np.pi, size=n)
        rx = radius * np.sin(theta) * np.cos(phi)
        ry = radius * np.sin(theta) * np.sin(phi)
        rz = radius * np.cos(theta)
        return rx, ry, rz

    def set_vel(self, pot):
        n = self.n
        rng = self.rng
        rnd = _von_neumann(rng, n)
        velocity = np.sqrt(-2 * rnd * pot)
        theta = np.arccos(rng.uniform(-1.0, 1.0, size=n))
        phi = rng.uniform(0.0, 2.0 * np.pi, size=n)
        vx = velocity * np.sin(theta) * np.cos(phi)
        vy = velocity * np.sin(theta) * np.sin(phi)
        vz = velocity * np.cos(theta)
        return vx, vy, vz

    def set_bodies(self):
        n = self.n
        ps = self.ps
        dev = ps.device
        ilist = np.arange(n)
        ps.id[...] = torch.as_tensor(ilist, device=dev)

        srand = self.rng.get_state()
        mass = np.asarray(self.imf.sample(n, self.rng), dtype=np.float64)
        mass = mass / float(np.sum(mass))
        eps2 = self.set_eps2(mass)
        self.rng.set_state(srand)

        rx, ry, rz = self.set_pos(self.rng.permutation(ilist))
        def put(name, v):
            getattr(ps, name)[...] = torch.as_tensor(
                v, dtype=getattr(ps, name).dtype, device=dev)
        put("mass", mass)
        put("eps2", eps2)
        put("rx", rx)
        put("ry", ry)
        put("rz", rz)

        ps.set_phi(ps)
        pot = ps.phi.detach().cpu().numpy().astype(np.float64)

        vx, vy, vz = self.set_vel(pot)
        put("vx", vx)
        put("vy", vy)
        put("vz", vz)

    def make_model(self):
        self.set_bodies()
        self.ps.com_to_origin()
        self.ps.to_nbody_units()
        self.ps.scale_to_virial()


def make_plummer(n, eps, imf, seed=None, mfrac=0.999, softening_type=0,
                 device=None):
    """Plummer sphere of n particles (plummer.py:213-221).  ``imf`` is a
    tuple (name, *args) naming an ``IMF`` class method, e.g.
    ("equalmass",), or an IMF sampler object."""
    if n < 2:
        n = 2
    if isinstance(imf, (tuple, list)):
        imf = getattr(IMF, imf[0])(*imf[1:])
    p = Plummer(n, eps, imf, seed=seed, mfrac=mfrac,
                softening_type=softening_type, device=device)
    p.make_model()
    return p.ps
