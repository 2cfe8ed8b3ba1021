"""Workload definitions: the benchmark configurations of the undulator step.

Config 0/1: planar undulator aw=1.7 (rms), lambda_u=15 mm, four modules of
125 periods with 0.5 m FODO drift sections (0.08 m quadrupoles,
k1=+-2/m^2); beam gamma 11357.8, rel. spread 1e-4, normalised emittance
0.4 um, beta 15 m, 3 kA; field 151x151, 0.5 mm half width; resonant
wavelength lambda_u(1+aw^2)/(2 gamma^2); one step per period.
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

# ---------------------------------------------------------------- configs
CONFIG1 = dict(name="config1", nslice=1024, npart=8192, nbins=4, ngrid=NGRID, gridmax=GRIDMAX,
               seed=7, shotnoise=True, longrange=True, sc_nz=1, sc_nphi=0, sc_ngrid=100,
               seed_power=0.0, timerun=True)
CONFIG0 = dict(name="config0", nslice=1, npart=8192, nbins=4, ngrid=NGRID, gridmax=GRIDMAX,
               seed=1, shotnoise=False, longrange=False, sc_nz=0, sc_nphi=0, sc_ngrid=100,
               seed_power=5e3, timerun=False)


CONFIG2 = dict(CONFIG1, name="config2", precision="fp32")


def config3_profile(ns=4096, peak=4000.0, sigma=1024.0, nmax=16384, nmin=256, nbins=4):
    """Uneven beam of config 3: I_s = peak exp(-((s - (ns-1)/2) / sigma)^2 / 2) and
    n_s = nmax I_s / peak macroparticles, clamped to [nmin, nmax] and rounded to
    a multiple of nbins (the mirrored beamlets)."""
    c = (ns - 1) / 2.0
    cur = [peak * math.exp(-0.5 * ((i - c) / sigma) ** 2) for i in range(ns)]
    cnt = [min(nmax, max(nmin, nbins * int(round(nmax * I / peak / nbins)))) for I in cur]
    return cur, cnt


CONFIG3 = dict(name="config3", nslice=4096, npart=None, nbins=4, ngrid=255, gridmax=GRIDMAX,
               seed=7, shotnoise=True, longrange=True, sc_nz=1, sc_nphi=0, sc_ngrid=100,
               seed_power=0.0, timerun=True, precision="fp32", uneven=True)
for _c in (CONFIG0, CONFIG1):
    _c["precision"] = "fp64"
CONFIGS = {c["name"]: c for c in (CONFIG0, CONFIG1, CONFIG2, CONFIG3)}


def counts_and_currents(cfg):
    if cfg.get("uneven"):
        cur, cnt = config3_profile(cfg["nslice"], nbins=cfg["nbins"])
        return cnt, cur
    return [cfg["npart"]] * cfg["nslice"], [CURRENT] * cfg["nslice"]


def describe(cfg):
    lsc = []
    if cfg["longrange"]:
        lsc.append("long-range LSC")
    if cfg["sc_nz"] > 0:
        lsc.append(f"short-range LSC nz={cfg['sc_nz']} nphi={cfg['sc_nphi']} "
                   f"{cfg['sc_ngrid']} radial cells")
    mode = "time-dependent SASE" if cfg["timerun"] else "steady state"
    prec = "FP64 plain form" if cfg.get("precision", "fp64") == "fp64" else \
        "mixed FP32 form (FP64 slice phase base, FP64 ADI residual correction)"
    if cfg.get("uneven"):
        cnt, _ = counts_and_currents(cfg)
        parts = (f"{cfg['nslice']} slices with a Gaussian current (peak 4 kA, sigma 1024 slices), "
                 f"{min(cnt)}..{max(cnt)} macroparticles per slice, {sum(cnt)} in total")
    else:
        parts = f"{cfg['nslice']} slices x {cfg['npart']} macroparticles"
    return (f"{cfg['name']}: {mode} {prec}, {parts} "
            f"(nbins {cfg['nbins']}), {cfg['ngrid']}x{cfg['ngrid']} field, "
            f"planar aw={AW} lambda_u={LAMBDAU} m, 4x125 periods + FODO, "
            + (", ".join(lsc) if lsc else "no space charge"))
