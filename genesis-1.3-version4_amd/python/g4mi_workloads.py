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


# The reference's Example1 deck (examples/Example1-SteadyState/Example1.in
# and Example1.lat) as data: steady state, helical aw = 0.84853, 255^2,
# 8192 macroparticles (nbins 8), 1104 steps with a record after each.
EXAMPLE1_IN = """&setup
rootname={root}
lattice={lat}
beamline=FEL
lambda0=1e-10
gamma0=11357.82
delz=0.045000
shotnoise=0
nbins = 8
&end

&lattice
zmatch=9.5
&end

&field
power=5e3
dgrid=2.000000e-04
ngrid=255
waist_size=30e-6
&end

&beam
current=3000
delgam=1.000000
ex=4.000000e-07
ey=4.000000e-07
&end

&track
&end
"""

EXAMPLE1_LAT = """D1: DRIFT = {{ l = 0.44}};
D2: DRIFT = {{ l = 0.24}};
QF: QUADRUPOLE = {{ l = 0.080000, k1= 2.000000 }};
QD: QUADRUPOLE = {{ l = 0.080000, k1= -2.000000 }};
UND: UNDULATOR = {{ lambdau=0.015000, nwig=266, aw=0.84853, helical= True{extra}}};

FODO: LINE={{UND,D1,QF,D2,UND,D1,QD,D2}};

FEL: LINE={{6*FODO}};
"""


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
        "mixed FP32 form (FP64 slice phase base, FP64 sums)"
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


# ---------------------------------------------------------------- runs
def build_workload(cfg):
    """Host side of a configuration: CSR offsets, the quiet-loaded beam (with
    shot noise for time-dependent runs), the initial field, the currents and
    the unrolled lattice."""
    import numpy as np
    import g4mi
    ns = cfg["nslice"]
    counts, currents = counts_and_currents(cfg)
    off = np.zeros(ns + 1, dtype=np.int64)
    off[1:] = np.cumsum(counts)
    slices = [slice_beam(current=c) for c in currents]
    beam = g4mi.load_beam(off, slices, cfg["nbins"], cfg["shotnoise"], cfg["seed"], LAMBDA_R,
                          1.0, cfg["timerun"])
    steps = g4mi.parse_lattice(lattice_text(), "FEL", DELZ, LAMBDA_R, GAMMA0)
    ng = cfg["ngrid"]
    field = np.zeros((ns, ng * ng, 2))
    if cfg["seed_power"] > 0:
        field[:] = g4mi.load_gauss(ng, cfg["gridmax"], LAMBDA_R, cfg["seed_power"], 30e-6)
    current = np.asarray(currents, dtype=np.float64)
    return off, beam, field, current, steps


def make_context(cfg, off, beam, field, current, precision):
    import g4mi
    ctx = g4mi.Context(off, cfg["ngrid"], cfg["gridmax"], LAMBDA_R, 1.0, cfg["timerun"], False,
                       precision)
    ctx.set_efield(cfg["longrange"], cfg["sc_nz"], cfg["sc_nphi"], cfg["sc_ngrid"], 0.0)
    ctx.upload_beam(beam, current)
    ctx.upload_field(field, 0)
    return ctx


def run_lattice(ctx, steps, warmup=0, timed=0, breakdown=0, keys=("power", "bunching")):
    """Runs every step of `steps` with a record after each: `warmup` steps,
    then `timed` steps bracketed by stream synchronisation, then the rest.
    Per-kernel-group times come from HIP events: over the timed steps in the
    FP64 form; in the mixed form (whose timed steps replay captured graphs)
    over `breakdown` steps launched directly between the warm-up and the
    timed steps.  Returns the seconds of the timed steps, the kernel times
    and the records [len(steps)][ns]."""
    import time
    import numpy as np
    import g4mi
    nst = len(steps)
    recs = {k: np.zeros((nst, ctx.ns)) for k in keys}
    L = g4mi.lib()

    def seg(a, b):
        if b <= a:
            return
        rows = (g4mi.DiagOut * (b - a))()
        for i in range(a, b):
            rows[i - a] = g4mi.DiagOut(*[recs[k][i].ctypes.data_as(g4mi._dp) if k in recs else None
                                         for k in g4mi.DIAG_FIELDS])
        arr = (g4mi.StepParams * (b - a))(*steps[a:b])
        done = g4mi._chk(L.g4mi_run(ctx.h, arr, b - a, GAMMA0, rows), "g4mi_run")
        assert done == b - a, (done, a, b)

    fp64 = ctx.precision == g4mi.FP64
    seg(0, warmup)
    ctx.sync()
    nb = 0
    if not fp64 and breakdown:
        nb = min(breakdown, nst - warmup - timed)
        ctx.set_timing(True)
        seg(warmup, warmup + nb)
        ctx.sync()
        kt = ctx.kernel_times()
        ctx.set_timing(False)
    elif fp64:
        ctx.set_timing(True)
    ctx.sync()
    t0 = time.perf_counter()
    seg(warmup + nb, warmup + nb + timed)
    ctx.sync()
    t1 = time.perf_counter()
    if fp64 or not nb:
        kt = ctx.kernel_times()
        ctx.set_timing(False)
    seg(warmup + nb + timed, nst)
    ctx.sync()
    return dict(seconds=t1 - t0, kt=kt, recs=recs)


def far_field_spectrum(field, first):
    """|FFT over the time window of the on-axis far-field amplitude|^2: the
    far field at zero angle is the sum of the near field over the grid
    (DiagField's farfield, Diagnostic.cpp:859-861), per beam slice i in
    window order (storage slice (i + first) mod ns)."""
    import numpy as np
    ns = field.shape[0]
    ff = field[..., 0].sum(axis=1) + 1j * field[..., 1].sum(axis=1)
    ff = np.roll(ff, -first)
    return np.abs(np.fft.fft(ff)) ** 2


def form_errors(r_ref, r, f_ref, first_ref, f, first, floor=1e-6):
    """Errors of a run against a reference run of the same workload:
    slice-summed P(z) and slice-averaged |b(z)| (max relative error over the
    steps where the reference's P is above `floor` of its maximum, and at
    the end), the final field (relative L2 and max-normalised) and the
    far-field spectrum (relative L2)."""
    import numpy as np
    assert first == first_ref, (first, first_ref)
    p0, p1 = r_ref["power"].sum(axis=1), r["power"].sum(axis=1)
    mask = p0 > floor * p0.max()
    b0, b1 = r_ref["bunching"].mean(axis=1), r["bunching"].mean(axis=1)
    d = f - f_ref
    s0, s1 = far_field_spectrum(f_ref, first_ref), far_field_spectrum(f, first)
    return dict(
        steps=int(len(p0)),
        p_rel_max=float(np.max(np.abs(p1[mask] / p0[mask] - 1))),
        p_rel_end=float(abs(p1[-1] / p0[-1] - 1)),
        b_rel_max=float(np.max(np.abs(b1 / b0 - 1))),
        b_rel_end=float(abs(b1[-1] / b0[-1] - 1)),
        field_rel_l2=float(np.linalg.norm(d) / np.linalg.norm(f_ref)),
        field_rel_max=float(np.max(np.abs(d)) / np.max(np.abs(f_ref))),
        spectrum_rel_l2=float(np.linalg.norm(s1 - s0) / np.linalg.norm(s0)),
        p_end_w=float(p0[-1]))
