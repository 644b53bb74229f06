"""Sakura integration of an ensemble of independent systems: the step of
tupan/integrator/sakura.py:22-50 (drift tau/2, pair drift with flag -1 over
tau/2, apply, the same with flag +1, drift tau/2) applied to every system at
once, each pair-drift stage one kernel launch over all systems, the drifts
and the application of dr, dv elementwise on the device arrays.

``evolve_graph`` records a block of steps once as a HIP graph (through
torch.cuda.CUDAGraph; every launch of the block, library kernels and
elementwise updates alike, becomes one graph node) and replays it, checking
the solver status once per replay instead of once per stage.  The replayed
steps run the same kernels in the same order as ``step``, so the results are
the same numbers."""
import torch


class SakuraEnsemble(object):
    def __init__(self, ps, ensemble):
        self.ps = ps
        self.ens = ensemble
        self.time = 0.0
        self.nsteps = 0

    def step(self, tau, check=True):
        self._advance(tau, check)
        self.time += tau
        self.nsteps += 1

    def _advance(self, tau, check):
        ps = self.ps
        ens = self.ens
        ps.rx += ps.vx * tau / 2
        ps.ry += ps.vy * tau / 2
        ps.rz += ps.vz * tau / 2
        for flag in (-1, 1):
            drx, dry, drz, dvx, dvy, dvz = ens.sakura(ps, tau / 2, flag,
                                                      check=check)
            ps.rx += drx
            ps.ry += dry
            ps.rz += drz
            ps.vx += dvx
            ps.vy += dvy
            ps.vz += dvz
        ps.rx += ps.vx * tau / 2
        ps.ry += ps.vy * tau / 2
        ps.rz += ps.vz * tau / 2

    def evolve(self, t_end, tau):
        while self.time < t_end - 0.5 * abs(tau):
            self.step(tau)
        return self.ps

    def capture(self, tau, nsteps):
        """HIP graph of `nsteps` steps of size tau on the particle arrays in
        place (their addresses are fixed in the graph)."""
        ens = self.ens
        ps = self.ps
        for k in ("drx", "dry", "drz", "dvx", "dvy", "dvz"):
            if k not in ps.__dict__:
                ps.register_auxiliary_attribute(k, "real")
        graph = torch.cuda.CUDAGraph()
        side = torch.cuda.Stream()
        side.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(side):
            with torch.cuda.graph(graph, stream=side):
                for _ in range(nsteps):
                    self._advance(tau, check=False)
        torch.cuda.current_stream().wait_stream(side)
        del ens
        return graph

    def evolve_graph(self, t_end, tau, steps_per_graph=8):
        """evolve() by replaying captured blocks of steps_per_graph steps
        (a shorter tail block is captured when needed)."""
        graphs = {}
        self.ens.reset_status()
        while self.time < t_end - 0.5 * abs(tau):
            left = int(round((t_end - self.time) / tau))
            k = max(1, min(steps_per_graph, left))
            if k not in graphs:
                graphs[k] = self.capture(tau, k)
            graphs[k].replay()
            for _ in range(k):
                self.time += tau
                self.nsteps += 1
        self.ens.check_status()
        return self.ps

    def energies(self):
        """(E per system, as an FP64 device tensor)."""
        ke, pe = self.ens.energies(self.ps)
        return ke + pe

    def max_rel_energy_error(self, e0):
        e1 = self.energies()
        return float(torch.max(torch.abs((e1 - e0) / e0)))
