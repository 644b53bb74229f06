"""Sakura integrator (tupan/integrator/sakura.py:22-159).

One step: drift tau/2; pairwise Kepler drift with flag -1 over tau/2 and
apply the summed changes; the same with flag +1; drift tau/2
(sakura.py:22-50).  'asakura' picks the shared step from Tstep as
eta / sqrt(1 + max_i(sum_j w2 - max_j w2)) and quantises it to the block
step (sakura.py:100-117).
"""
from . import Base
from ..lib import extensions


def sakura_step(ps, tau):
    ps.rx += ps.vx * tau / 2
    ps.ry += ps.vy * tau / 2
    ps.rz += ps.vz * tau / 2

    extensions.sakura.calc(ps, ps, tau / 2, -1)
    ps.rx += ps.drx
    ps.ry += ps.dry
    ps.rz += ps.drz
    ps.vx += ps.dvx
    ps.vy += ps.dvy
    ps.vz += ps.dvz

    extensions.sakura.calc(ps, ps, tau / 2, 1)
    ps.rx += ps.drx
    ps.ry += ps.dry
    ps.rz += ps.drz
    ps.vx += ps.dvx
    ps.vy += ps.dvy
    ps.vz += ps.dvz

    ps.rx += ps.vx * tau / 2
    ps.ry += ps.vy * tau / 2
    ps.rz += ps.vz * tau / 2
    return ps


class Sakura(Base):
    PROVIDED_METHODS = ["sakura", "asakura"]

    def __init__(self, eta, time, ps, method, **kwargs):
        super(Sakura, self).__init__(eta, time, ps, **kwargs)
        self.method = method

    def get_sakura_tstep(self, ps, eta, tau):
        ps.set_tstep(ps, eta)
        iw2_a = (eta / ps.tstep)**2
        iw2_b = (eta / ps.tstepij)**2
        diw2 = iw2_a - iw2_b
        w2_sakura = float(diw2.max())
        dt_sakura = eta / (1 + w2_sakura)**0.5
        ps.tstep[...] = dt_sakura
        return self.get_min_block_tstep(ps, tau)

    def do_step(self, ps, tau):
        if "asakura" in self.method:
            tau = self.get_sakura_tstep(ps, self.eta, tau)
        ps = sakura_step(ps, tau)
        type(ps).t_curr += tau
        ps.tstep[...] = tau
        ps.time += tau
        ps.nstep += 1
        return ps
