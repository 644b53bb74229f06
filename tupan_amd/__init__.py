"""tupan_amd: MI355X (gfx950) implementation of tupan's all-pairs gravity
kernels (Phi, Acc, AccJerk, SnapCrackle, Tstep), the Sakura two-body drift
and the Kepler two-body solver, behind tupan's extension-kernel API
(tupan.lib.extensions), with device-resident particle systems and the
Hermite, SIA and Sakura integrators that call them.
"""
__version__ = "0.1.0"
