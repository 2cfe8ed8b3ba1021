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
