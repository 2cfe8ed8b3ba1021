"""Workload definitions shared by the tests, the smoke entry and bench.py.

Config 0/1 of the undulator-step benchmark: planar undulator aw=1.7 (rms),
lambda_u=15 mm, four modules of 125 periods with 0.5 m FODO drift sections
(0.08 m quadrupoles, k1=+-2/m^2); beam gamma 11357.8, rel. spread 1e-4,
normalised emittance 0.4 um, beta 15 m, 3 kA; field 151x151, 0.5 mm half
width; resonant wavelength lambda_u(1+aw^2)/(2 gamma^2); one step per
period.
"""
import math

GAMMA0 = 11357.8
AW = 1.7
LAMBDAU = 0.015
LAMBDA_R = LAMBDAU * (1 + AW * AW) / (2 * GAMMA0 * GAMMA0)
DELZ = 0.015
CURRENT = 3000.0
NGRID = 151
GRIDMAX = 0.5e-3


def slice_beam(current=CURRENT, gamma=GAMMA0):
    return dict(gamma=gamma, delgam=1e-4 * gamma, current=current, ex=0.4e-6, ey=0.4e-6,
                betax=15.0, betay=15.0, alphax=0.0, alphay=0.0)


def fodo_elements(nund=4, nwig=125):
    """UND, D, Q, D, UND, ... (drift halves of (0.5-0.08)/2 m), no trailing drift."""
    els = []
    half = 0.5 * (0.5 - 0.08)
    for i in range(nund):
        els.append(("undulator", dict(aw=AW, lambdau=LAMBDAU, nwig=nwig)))
        if i < nund - 1:
            els.append(("drift", dict(l=half)))
            els.append(("quadrupole", dict(l=0.08, k1=2.0 if i % 2 == 0 else -2.0)))
            els.append(("drift", dict(l=half)))
    return els


def lattice_text(nund=4, nwig=125):
    """The same beamline as a lattice file for the library's parser."""
    half = 0.5 * (0.5 - 0.08)
    s = [f"UND: UNDULATOR = {{lambdau={LAMBDAU!r}, nwig={nwig}, aw={AW!r}}};",
         f"D: DRIFT = {{l={half!r}}};",
         "QF: QUADRUPOLE = {l=0.08, k1=2.0};",
         "QD: QUADRUPOLE = {l=0.08, k1=-2.0};"]
    items = []
    for i in range(nund):
        items.append("UND")
        if i < nund - 1:
            items += ["D", "QF" if i % 2 == 0 else "QD", "D"]
    s.append("FEL: LINE = {" + ", ".join(items) + "};")
    return "\n".join(s) + "\n"


def gaussian_current(ns, peak=4000.0, sigma_frac=0.25):
    """Gaussian current profile over ns slices (sigma = sigma_frac*ns)."""
    c = (ns - 1) / 2.0
    sig = sigma_frac * ns
    return [peak * math.exp(-0.5 * ((i - c) / sig) ** 2) for i in range(ns)]
