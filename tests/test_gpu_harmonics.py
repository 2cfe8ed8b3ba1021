"""GPU parity of harmonic fields (Field with harm > 1) against the oracle:
the push sums every field's coupling (BeamSolver.cpp:23-32, 63-79,
144-163), each odd harmonic is driven by the source at h theta
(FieldSolverADI.cpp:19-50) and propagated by its own ADI, and slippage
moves every field alike.  FP64 bar: beam and fields within 1e-12 after a
few steps, P(z) of every field within 1e-10 over a run."""
import ctypes as C
import sys

import numpy as np
import pytest

import cases
import g4mi
import harness as H
import oracle_py as O

pytestmark = pytest.mark.gpu

HARM = [(3, 41, 0.25e-3), (5, 31, 0.2e-3)]


def _log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _pair(ns, timerun, seed_power=5e7):
    sim = H.make_state([1024] * ns, ngrid=63, gridmax=0.3e-3, seed_power=seed_power, timerun=timerun,
                       threads=16)
    sim.theta[:] += 0.2 * np.sin(3 * sim.theta) + 0.1 * np.sin(5 * sim.theta)  # harmonic bunching
    sim.add_harmonics(HARM)
    rng = np.random.default_rng(1)
    for f in sim.hfields:
        f[:] = rng.normal(size=f.shape) * 1e4
    ctx = H.to_gpu(sim)
    for k, (h, ng, gm) in enumerate(HARM):
        assert ctx.add_harmonic(h, ng, gm) == k
        ctx.upload_harmonic(k, sim.hfields[k])
    return sim, ctx


def test_harmonic_fields_steps():
    _log("harmonics 3 and 5: 4 slices x 1024, 10 undulator steps, GPU vs oracle")
    st = H.steps()[:10]
    sim, ctx = _pair(4, timerun=False)
    ctx.run(H.gsteps(st), cases.GAMMA0, record=("power",))
    sim.run(st, cases.GAMMA0, record=("power",))
    b = ctx.download_beam()
    for k in ("gamma", "theta", "x", "y", "px", "py"):
        ref = getattr(sim, k)
        err = np.max(np.abs(b[k] - ref)) / np.max(np.abs(ref))
        assert err < 1e-12, (k, err)
    f, _ = ctx.download_field()
    assert np.max(np.abs(f - sim.field)) < 1e-12 * np.max(np.abs(sim.field))
    for k in range(len(HARM)):
        fh = ctx.download_harmonic(k)
        e = np.max(np.abs(fh - sim.hfields[k])) / np.max(np.abs(sim.hfields[k]))
        _log(f"  harmonic {HARM[k][0]}: field max diff / max |E| = {e:.3e}")
        assert e < 1e-12
        np.testing.assert_allclose(ctx.harmonic_power(k), sim.hpower(k), rtol=1e-12)
    ctx.close()


def test_harmonic_sase_end_to_end():
    """Time-dependent, 12 slices with shot noise and slippage, 150 steps:
    fundamental P(z) and the harmonic powers at the end within 1e-10."""
    _log("harmonics 3 and 5: SASE, 12 slices, 150 steps, GPU vs oracle")
    st = H.steps()[:150]
    sim, ctx = _pair(12, timerun=True, seed_power=0.0)
    g = ctx.run(H.gsteps(st), cases.GAMMA0, record=("power", "bunching"))
    r = sim.run(st, cases.GAMMA0, record=("power", "bunching"))
    ep = H.rel_err(g["power"][1:], r["power"][1:])
    eb = H.rel_err(g["bunching"], r["bunching"])
    eh = [H.rel_err(ctx.harmonic_power(k), sim.hpower(k)) for k in range(len(HARM))]
    _log(f"  P(z) rel err {ep:.3e}  |b(z)| {eb:.3e}  harmonic power {eh}")
    assert ep < 1e-10 and eb < 1e-10 and max(eh) < 1e-10
    f, first = ctx.download_field()
    assert first == sim.s.first
    ctx.close()


def test_harmonic_api_limits():
    sim = H.make_state([256] * 2, ngrid=31, gridmax=0.3e-3)
    ctx = H.to_gpu(sim)
    ctx.add_harmonic(3, 31, 0.3e-3)
    with pytest.raises(g4mi.G4miError):
        ctx.add_harmonic(3, 31, 0.3e-3)  # exists already
    with pytest.raises(g4mi.G4miError):
        ctx.set_field_solver(fft=True)  # the spectral solver runs the fundamental only
    ctx.close()
    mix = H.to_gpu(sim, precision=g4mi.FP32)
    with pytest.raises(g4mi.G4miError):
        mix.add_harmonic(3, 31, 0.3e-3)
