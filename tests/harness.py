"""Builds matched oracle/GPU states for parity tests (test infrastructure)."""
import ctypes as C

import numpy as np

import cases
import g4mi
import oracle_py as O


def make_state(counts, ngrid=cases.NGRID, gridmax=cases.GRIDMAX, seed_power=0.0,
               shotnoise_seed=7, timerun=True, currents=None, sc=(0, 0, 100, 0.0),
               longrange=False, threads=8, nbins=4):
    """Load a beam (and a Gaussian seed if seed_power > 0) with the
    oracle's loader; returns the oracle Sim."""
    ns = len(counts)
    sim = O.Sim(counts, ngrid, gridmax, cases.LAMBDA_R, timerun=timerun, sc=sc,
                longrange=longrange, threads=threads)
    cur = currents if currents is not None else [cases.CURRENT] * ns
    O.load_beam(sim, [cases.slice_beam(current=c) for c in cur], None, nbins,
                shotnoise_seed if timerun else None, lambda0=cases.LAMBDA_R)
    if seed_power > 0:
        O.load_gauss_all(sim, seed_power, 30e-6, cases.LAMBDA_R)
    return sim


def to_gpu(sim, precision=g4mi.FP64):
    s = sim.s
    ctx = g4mi.Context(sim.off, s.ngrid, s.gridmax, cases.LAMBDA_R, sample=s.sample,
                       timerun=bool(s.timerun), periodic=bool(s.periodic), precision=precision)
    ctx.set_efield(longrange=bool(s.longrange), nz=s.sc_nz, nphi=s.sc_nphi, ngrid_r=s.sc_ngrid,
                   rmax=s.sc_rmax)
    ctx.upload_beam({k: getattr(sim, k) for k in ("gamma", "theta", "x", "y", "px", "py")},
                    sim.current)
    ctx.upload_field(sim.field, s.first)
    return ctx


def steps(nund=4, nwig=125):
    return O.unroll(cases.fodo_elements(nund, nwig), cases.DELZ, cases.LAMBDA_R, cases.GAMMA0)


def gsteps(st):
    return g4mi.steps_from(st)


def rel_err(a, b, floor_frac=1e-12):
    """max |a-b| / max(|b|, floor), floor = floor_frac * max|b| over the run."""
    a, b = np.asarray(a, float), np.asarray(b, float)
    floor = floor_frac * np.max(np.abs(b)) if b.size else 0.0
    den = np.maximum(np.abs(b), floor)
    den[den == 0] = 1.0
    return float(np.max(np.abs(a - b) / den)) if b.size else 0.0


def oracle_call(name, sim, *args):
    getattr(O.lib(), name)(C.byref(sim.s), *args)


def aux_reference(beam, field, first, off, cur, ngrid, dgrid, lambda0):
    """Textbook numpy reductions of a downloaded state for the auxiliary and
    global diagnostics (DiagBeam, src/Core/Diagnostic.cpp:415-445, 505-519,
    552-607; DiagField :755-861, 900-969)."""
    ns = len(off) - 1
    out = {k: np.zeros(ns) for k in ("pxposition", "pyposition", "xmin", "xmax", "pxmin", "pxmax",
                                      "ymin", "ymax", "pymin", "pymax", "emin", "emax", "nf_phase",
                                      "ff_phase")}
    bounds = {"x": (1e5, -1e5), "px": (1e5, -1e5), "y": (1e5, -1e5), "py": (1e5, -1e5),
              "gamma": (1e7, 1.0)}
    names = {"x": "x", "px": "px", "y": "y", "py": "py", "gamma": "e"}
    gw = dict(c=0.0, g1=0.0, g2=0.0, x1=0.0, x2=0.0, y1=0.0, y2=0.0)
    cen = (ngrid // 2) * ngrid + ngrid // 2
    shift = -0.5 * (ngrid - 1)
    iy, ix = np.divmod(np.arange(ngrid * ngrid), ngrid)
    dx, dy = ix + shift, iy + shift
    fw = dict(p=0.0, x1=0.0, x2=0.0, y1=0.0, y2=0.0, ff=0.0, inten=0.0)
    for i in range(ns):
        a, b = int(off[i]), int(off[i + 1])
        n = b - a
        for k in ("px", "py"):
            out[k + "position"][i] = beam[k][a:b].sum() / n if n else 0.0
        for k, (lo, hi) in bounds.items():
            v = beam[k][a:b]
            out[names[k] + "min"][i] = min(lo, v.min()) if n else lo
            out[names[k] + "max"][i] = max(hi, v.max()) if n else hi
        fs = (i + first) % ns
        e = field[fs, :, 0] + 1j * field[fs, :, 1]
        out["nf_phase"][i] = np.arctan2(e[cen].imag, e[cen].real)
        ffs = e.sum()
        out["ff_phase"][i] = np.arctan2(ffs.imag, ffs.real)
        w = np.abs(e) ** 2
        fw["p"] += w.sum()
        fw["x1"] += (dx * w).sum()
        fw["x2"] += (dx * dx * w).sum()
        fw["y1"] += (dy * w).sum()
        fw["y2"] += (dy * dy * w).sum()
        fw["ff"] += abs(ffs) ** 2 * dgrid * dgrid
        fw["inten"] += w[cen]
        g = beam["gamma"][a:b]
        gw["c"] += cur[i]
        gw["g1"] += cur[i] * g.mean()
        gw["g2"] += cur[i] * (g * g).mean()
        gw["x1"] += cur[i] * beam["x"][a:b].mean()
        gw["x2"] += cur[i] * (beam["x"][a:b] ** 2).mean()
        gw["y1"] += cur[i] * beam["y"][a:b].mean()
        gw["y2"] += cur[i] * (beam["y"][a:b] ** 2).mean()
    c = gw["c"]
    g1, g2 = gw["g1"] / c, gw["g2"] / c
    x1, x2, y1, y2 = gw["x1"] / c, gw["x2"] / c, gw["y1"] / c, gw["y2"] / c
    out.update(g_energy=g1, g_energyspread=np.sqrt(abs(g2 - g1 * g1)), g_xposition=x1,
               g_xsize=np.sqrt(abs(x2 - x1 * x1)), g_yposition=y1, g_ysize=np.sqrt(abs(y2 - y1 * y1)))
    p = fw["p"]
    fx1, fx2, fy1, fy2 = fw["x1"] / p, fw["x2"] / p, fw["y1"] / p, fw["y2"] / p
    ks = 2 * np.pi / lambda0
    eev, vacimp = 510998.95069, 376.73
    out.update(g_f_xposition=fx1 * dgrid, g_f_xsize=np.sqrt(abs(fx2 - fx1 * fx1)) * dgrid,
               g_f_yposition=fy1 * dgrid, g_f_ysize=np.sqrt(abs(fy2 - fy1 * fy1)) * dgrid,
               g_nf_intensity=fw["inten"] / ns * eev * eev / ks / ks / vacimp,
               g_ff_intensity=fw["ff"] / ns)
    return out


def check_aux(d, ref, beam, tol):
    """Compare the auxiliary/global diagnostics with aux_reference.  Means
    that the quiet loader makes zero (positions, momenta, field centroids)
    are rounding noise, so they are compared against the matching spread;
    phases in radians; everything else relative."""
    spread = {"pxposition": np.std(beam["px"]), "pyposition": np.std(beam["py"]),
              "g_xposition": ref["g_xsize"], "g_yposition": ref["g_ysize"],
              "g_f_xposition": ref["g_f_xsize"], "g_f_yposition": ref["g_f_ysize"]}
    for k, sc in spread.items():
        e = float(np.max(np.abs(np.atleast_1d(d[k]) - np.atleast_1d(ref[k])))) / sc
        assert e < tol, (k, e)
    for k in ("nf_phase", "ff_phase"):
        e = float(np.max(np.abs(np.angle(np.exp(1j * (d[k] - ref[k]))))))
        assert e < tol, (k, e)
    for k in ("g_energy", "g_xsize", "g_ysize", "g_f_xsize", "g_f_ysize", "g_nf_intensity",
              "g_ff_intensity"):
        e = rel_err(np.atleast_1d(d[k]), np.atleast_1d(ref[k]), 1e-12)
        assert e < tol, (k, e)
    for k in ("xmin", "xmax", "pxmin", "pxmax", "ymin", "ymax", "pymin", "pymax", "emin", "emax"):
        assert np.array_equal(d[k], ref[k]), k
