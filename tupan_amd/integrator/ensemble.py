"""Sakura integration of an ensemble of independent systems: the step of
tupan/integrator/sakura.py:22-50 (drift tau/2, pair drift with flag -1 over
tau/2, apply, the same with flag +1, drift tau/2) applied to every system at
once, each pair-drift stage one kernel launch over all systems, the drifts
and the application of dr, dv elementwise on the device arrays."""
import torch


class SakuraEnsemble(object):
    def __init__(self, ps, ensemble):
        self.ps = ps
        self.ens = ensemble
        self.time = 0.0
        self.nsteps = 0

    def step(self, tau):
        ps = self.ps
        ens = self.ens
        ps.rx += ps.vx * tau / 2
        ps.ry += ps.vy * tau / 2
        ps.rz += ps.vz * tau / 2
        for flag in (-1, 1):
            drx, dry, drz, dvx, dvy, dvz = ens.sakura(ps, tau / 2, flag)
            ps.rx += drx
            ps.ry += dry
            ps.rz += drz
            ps.vx += dvx
            ps.vy += dvy
            ps.vz += dvz
        ps.rx += ps.vx * tau / 2
        ps.ry += ps.vy * tau / 2
        ps.rz += ps.vz * tau / 2
        self.time += tau
        self.nsteps += 1

    def evolve(self, t_end, tau):
        while self.time < t_end - 0.5 * abs(tau):
            self.step(tau)
        return self.ps

    def energies(self):
        """(E per system, as an FP64 device tensor)."""
        ke, pe = self.ens.energies(self.ps)
        return ke + pe

    def max_rel_energy_error(self, e0):
        e1 = self.energies()
        return float(torch.max(torch.abs((e1 - e0) / e0)))
