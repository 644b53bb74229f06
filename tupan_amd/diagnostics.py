"""Energy and momentum diagnostics (tupan/simulation.py:44-129).

``Diagnostic`` keeps the reference's constructor, header and report line:
time, dtime, ke, pe, te, virial, eerr = (E - E0)/(-PE) (simulation.py:109,
normalised by -PE), geerr (running RMS of eerr, :111-112), |dcom_r|,
|dcom_v|, |dL|, |dJ| and wall-clock time, in the reference's column widths,
so report files compare line by line (the wall-clock column aside).  ``energy`` gives E = K + U with U from FP64 Phi and exactly
rounded sums.
"""
import math
import sys
import time


def energy(ps):
    return ps.kinetic_energy + ps.potential_energy


def myprint(data, fname, fmode):
    """simulation.py:30-37: '<stdout>', '<stderr>' or a file name."""
    if fname == "<stdout>":
        print(data, file=sys.stdout)
    elif fname == "<stderr>":
        print(data, file=sys.stderr)
    else:
        with open(fname, fmode) as fobj:
            print(data, file=fobj)


class Diagnostic(object):
    """simulation.py:40-129: Diagnostic(fname, time, ps, report_freq=4,
    pn_order=0).  The first report writes the header (mode 'w'), every
    report_freq-th call to diagnostic_report appends one line; every report
    is also kept in ``rows`` (with |dE/E| added).  fname None keeps the rows
    only."""
    HEADER = ("#00:time", "#01:dtime", "#02:ke", "#03:pe", "#04:te",
              "#05:virial", "#06:eerr", "#07:geerr", "#08:com_r",
              "#09:com_v", "#10:lmom", "#11:amom", "#12:wct")
    HEADER_FMT = ("{0:13s} {1:10s} {2:10s} {3:10s} {4:15s} {5:10s} {6:10s} "
                  "{7:10s} {8:10s} {9:10s} {10:10s} {11:10s} {12:13s}")
    ROW_FMT = ("{time:< 13.6e} {dtime:< 10.3e} {ke:< 10.3e} {pe:< 10.3e} "
               "{te:< 15.8e} {virial:< 10.3e} {eerr:< 10.3e} "
               "{geerr:< 10.3e} {com_r:< 10.3e} {com_v:< 10.3e} "
               "{lmom:< 10.3e} {amom:< 10.3e} {wct:< 13.6e}")

    def __init__(self, fname=None, time=0.0, ps=None, report_freq=4,
                 pn_order=0):
        self.fname = fname
        self.time = time
        self.report_freq = report_freq
        self.include_pn_corrections = bool(pn_order)
        self.nreport = 0
        self.is_initialized = False
        self.rows = []

    def initialize(self, ps):
        self.ke0 = ps.kinetic_energy
        self.pe0 = ps.potential_energy
        self.te0 = self.ke0 + self.pe0
        self.com_r0 = ps.com_r
        self.com_v0 = ps.com_v
        self.lmom0 = ps.linear_momentum
        self.amom0 = ps.angular_momentum
        self.count = 0
        self.ceerr = 0.0
        self.t0 = time.perf_counter()
        self.print_header()
        self.is_initialized = True

    def print_header(self):
        if self.fname is not None:
            myprint(self.HEADER_FMT.format(*self.HEADER), self.fname, "w")

    def diagnostic_report(self, ps):
        if not self.is_initialized:
            self.initialize(ps)
        t_curr = ps.t_curr
        if self.nreport % self.report_freq == 0:
            self.print_diagnostic(t_curr, t_curr - self.time, ps)
        self.nreport += 1

    def print_diagnostic(self, t, dtime, ps):
        self.time = t
        ke = ps.kinetic_energy
        pe = ps.potential_energy
        te = ke + pe
        virial = ps.virial_energy
        eerr = (te - self.te0) / (-pe)
        self.count += 1
        self.ceerr += eerr**2
        geerr = math.sqrt(self.ceerr / self.count)

        def dist(a, b):
            return float(((a - b)**2).sum())**0.5
        row = {"time": t, "dtime": dtime, "ke": ke, "pe": pe, "te": te,
               "virial": virial, "eerr": eerr, "geerr": geerr,
               "com_r": dist(ps.com_r, self.com_r0),
               "com_v": dist(ps.com_v, self.com_v0),
               "lmom": dist(ps.linear_momentum, self.lmom0),
               "amom": dist(ps.angular_momentum, self.amom0),
               "wct": time.perf_counter() - self.t0}
        if self.fname is not None:
            myprint(self.ROW_FMT.format(**row), self.fname, "a")
        row["dE_E"] = abs((te - self.te0) / self.te0)
        self.rows.append(row)
        return row
