"""Few-body leaves of the SIA hierarchy (tupan/integrator/fewbody.py:13-57):
a single particle drifts; a pair is advanced exactly by the Kepler solver.
With PN corrections a drift also moves pn_mr, and a pair raises
NotImplementedError, as in the reference: its Kepler solver has no PN
terms."""
from ..lib import extensions


class FewBody(object):
    @staticmethod
    def drift(ips, tau):
        ips.rx += ips.vx * tau
        ips.ry += ips.vy * tau
        ips.rz += ips.vz * tau
        if ips.include_pn_corrections:
            ips.pn_drift_com_r(tau)
        return ips

    @staticmethod
    def kepler_solver(ips, tau):
        if ips.include_pn_corrections:
            raise NotImplementedError(
                "The current version of the Kepler-solver does not include "
                "post-Newtonian corrections.")
        extensions.kepler.calc(ips, ips, tau)
        return ips

    @staticmethod
    def evolve(ips, tau):
        if ips.n == 0:
            return ips
        if ips.n == 1:
            return FewBody.drift(ips, tau)
        return FewBody.kepler_solver(ips, tau)
