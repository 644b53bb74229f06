"""Few-body leaves of the SIA hierarchy (tupan/integrator/fewbody.py:13-57):
a single particle drifts; a pair is advanced exactly by the Kepler solver."""
from ..lib import extensions


class FewBody(object):
    @staticmethod
    def drift(ips, tau):
        ips.rx += ips.vx * tau
        ips.ry += ips.vy * tau
        ips.rz += ips.vz * tau
        return ips

    @staticmethod
    def kepler_solver(ips, tau):
        extensions.kepler.calc(ips, ips, tau)
        return ips

    @staticmethod
    def evolve(ips, tau):
        if ips.n == 0:
            return ips
        if ips.n == 1:
            return FewBody.drift(ips, tau)
        return FewBody.kepler_solver(ips, tau)
