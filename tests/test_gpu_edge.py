"""Ragged input with empty slices through the FP64 step on the device.

Slices without particles occur in one4one runs (after sorting) and in
imported beams.  The reference handles them by skipping: Beam::getSize
returns 1 for a slice without current or particles (src/Core/Beam.cpp:42-67),
the diagnostics normalise by 1 (Diagnostic.cpp:447-451), the deposit adds
nothing.  The oracle follows the same code paths; the device must agree
with it on every non-empty slice and give zeros (not NaN) on the empty
ones."""
import numpy as np
import pytest

import cases
import harness as H
import oracle_py as O

pytestmark = pytest.mark.gpu


def _with_empty(sim0, pattern):
    """sim0's slices rearranged by `pattern` (old slice index, or None for an
    empty slice with zero current); every slice keeps a copy of a seeded
    field slice."""
    counts = [0 if j is None else int(sim0.off[j + 1] - sim0.off[j]) for j in pattern]
    sim = O.Sim(counts, sim0.s.ngrid, sim0.s.gridmax, cases.LAMBDA_R, timerun=True, sc=(1, 0, 100, 0.0),
                longrange=True, threads=8)
    for k in ("gamma", "theta", "x", "y", "px", "py"):
        src = getattr(sim0, k)
        getattr(sim, k)[:] = np.concatenate([src[sim0.off[j]:sim0.off[j + 1]] for j in pattern if j is not None])
    sim.current[:] = [0.0 if j is None else sim0.current[j] for j in pattern]
    for i, j in enumerate(pattern):
        sim.field[i] = sim0.field[0 if j is None else j]
    return sim


def test_empty_slices_end_to_end():
    """Five slices, the 2nd and the last empty, long- and short-range space
    charge, 40 undulator steps with slippage: P(z) of every slice and |b(z)|
    and the particles of the non-empty slices within 1e-10 of the oracle;
    nothing NaN."""
    print("FP64 step with empty slices (5 slices, 2 empty), 40 steps", flush=True)
    sim0 = H.make_state([512, 768, 256], seed_power=5e3, longrange=True, sc=(1, 0, 100, 0.0))
    sim = _with_empty(sim0, [0, None, 1, 2, None])
    ctx = H.to_gpu(sim)
    st = H.steps()[:40]
    g = ctx.run(H.gsteps(st), cases.GAMMA0, record=("power", "bunching"))
    r = sim.run(st, cases.GAMMA0, record=("power", "bunching"))
    assert np.all(np.isfinite(g["power"])) and np.all(np.isfinite(g["bunching"]))
    ep = H.rel_err(g["power"], r["power"])
    full = [0, 2, 3]
    eb = H.rel_err(g["bunching"][:, full], r["bunching"][:, full])
    assert np.all(g["bunching"][:, [1, 4]] == 0)
    b = ctx.download_beam()
    eg = max(H.rel_err(b[k], getattr(sim, k)) for k in ("gamma", "theta", "x", "y"))
    print(f"  P(z) rel err {ep:.3e}, |b(z)| {eb:.3e}, particles {eg:.3e}", flush=True)
    assert ep < 1e-10 and eb < 1e-10 and eg < 1e-10
    ctx.close()
