"""Initial mass functions (tupan/ics/imf.py:19-90, plus samplers the
reference does not offer).

``IMF.equalmass()`` and ``IMF.salpeter1955(mlow, mhigh)`` follow the
reference's ``IMFSample`` rejection sampler (imf.py:19-64): the IMF is
normalised by quad(imf(m)/m) over its full range, its peak found with
fminbound, and masses drawn log-uniformly and accepted under the scaled
curve, in batches, from the supplied RandomState.

The reference's Salpeter IMF only accepts mlow >= 0.4 (imf.py:84-90), so the
power-law mass function of the block-time-step workload (slope -2.35 on
[0.1, 10]) uses ``IMF.power_law``: inverse-transform sampling of
dN/dm ~ m^slope, one uniform draw per mass.  ``IMF.log_uniform`` draws
log(m) uniformly on [log mlow, log mhigh] (the ensemble workload).
"""
import numpy as np


class IMFSample(object):
    """Rejection sampler of the reference (imf.py:19-64)."""

    def __init__(self, imf_func, min_mlow, max_mhigh, mlow, mhigh):
        if mlow < min_mlow or mhigh > max_mhigh:
            raise ValueError("Mass range outer of limits"
                             " [{0}:{1}].".format(min_mlow, max_mhigh))
        from scipy.integrate import quad
        from scipy.optimize import fminbound
        norm, _ = quad(lambda m: imf_func(m) / m, min_mlow, max_mhigh)
        func = lambda m: imf_func(m) / norm  # noqa: E731
        mpeak = float(fminbound(lambda m: -func(m), min_mlow, max_mhigh,
                                xtol=1.0e-8))
        self.func = func
        self.peak = func(mpeak)
        self.mlow = mlow
        self.mpeak = mpeak
        self.mhigh = mhigh
        self.min_mlow = min_mlow
        self.max_mhigh = max_mhigh
        self._sample = None
        self._mtot = None
        self.equal = mlow == mhigh

    def sample(self, n, rng):
        size = n
        ran_mass = []
        while len(ran_mass) < n:
            ran_imf = rng.uniform(0.0, self.peak, size=size)
            ran_mass = np.exp(rng.uniform(np.log(self.mlow), np.log(self.mhigh),
                                          size=size))
            ran_mass = ran_mass[np.where(ran_imf < self.func(ran_mass))]
            size *= int(1 + n / max(len(ran_mass), 1))
        self._sample = ran_mass[:n]
        self._mtot = float(np.sum(self._sample))
        return self._sample


class PowerLaw(object):
    """dN/dm ~ m^slope on [mlow, mhigh], inverse-transform sampled."""

    def __init__(self, slope, mlow, mhigh):
        if not (0 < mlow < mhigh):
            raise ValueError("need 0 < mlow < mhigh")
        self.slope = float(slope)
        self.mlow = float(mlow)
        self.mhigh = float(mhigh)
        self.equal = False

    def sample(self, n, rng):
        u = rng.random_sample(n)
        k = self.slope + 1.0
        if abs(k) < 1e-12:
            return self.mlow * (self.mhigh / self.mlow)**u
        a = self.mlow**k
        b = self.mhigh**k
        return (a + u * (b - a))**(1.0 / k)


class LogUniform(object):
    """log m uniform on [log mlow, log mhigh]."""

    def __init__(self, mlow, mhigh):
        self.mlow = float(mlow)
        self.mhigh = float(mhigh)
        self.equal = mlow == mhigh

    def sample(self, n, rng):
        return np.exp(rng.uniform(np.log(self.mlow), np.log(self.mhigh),
                                  size=n))


class IMF(object):
    @classmethod
    def equalmass(cls):
        """imf.py:74-82: all masses equal."""
        return IMFSample(lambda m: (1.0 + m) - m, 0.1, 10.0, 1.0, 1.0)

    @classmethod
    def salpeter1955(cls, mlow, mhigh):
        """imf.py:84-90: dN/dlog m ~ m^-1.35 on [0.4, 120]."""
        return IMFSample(lambda m: m**(-1.35), 0.4, 120.0, mlow, mhigh)

    @classmethod
    def power_law(cls, slope, mlow, mhigh):
        return PowerLaw(slope, mlow, mhigh)

    @classmethod
    def log_uniform(cls, mlow, mhigh):
        return LogUniform(mlow, mhigh)
