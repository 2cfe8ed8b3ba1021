"""GPU tests of the learned field step (the FNO surrogate).

Forward output and parameter gradients of the library's BF16-MFMA network
against the float64 CPU reference (tests/fno_reference.py) on the same
parameters and inputs; training on (E, S, E_next) pairs captured from the
solver's own ADI steps, with the gain over the untrained network and over
the field-unchanged prediction; and the ADI-replacement mode with its
energy-balance fallback.  BF16 operands carry 8 significant bits, so the
network-level bounds are a few percent (stated per test)."""
import sys

import numpy as np
import pytest

import cases
import g4mi
import harness as H
import fno_reference as R

pytestmark = pytest.mark.gpu


def _log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


SMALL = dict(ngrid=33, width=8, modes=8, layers=2, hidden=16)


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def test_fno_forward_matches_reference():
    """Random parameters (scaled up from the default initialisation so every
    path contributes) and inputs: output within 3 % (relative L2) of the
    float64 reference."""
    sh = SMALL
    f = g4mi.Fno(max_batch=3, seed=3, **sh)
    rng = np.random.default_rng(5)
    p = f.get_params().astype(np.float64)
    for name, o, n in f.param_blocks():
        if name.startswith("R"):
            p[o:o + n] = rng.normal(size=n) * 0.3
    f.set_params(p.astype(np.float32))
    N = sh["ngrid"]
    inp = rng.normal(size=(3, N * N, 4)).astype(np.float32)
    out = f.forward(inp)
    _, _, ref = R.loss_and_grad(p, inp, np.zeros((3, N * N, 2)), N, sh["width"], sh["modes"],
                                sh["layers"], sh["hidden"])
    e = _rel(out, ref)
    _log(f"FNO forward vs float64 reference: rel L2 {e:.3e}")
    assert e < 3e-2
    f.close()


def test_fno_gradients_match_reference():
    """Loss and every parameter block's gradient against torch autograd in
    float64: loss within 2 %, each block's gradient within 5 % (relative L2)
    with cosine similarity above 0.998."""
    sh = SMALL
    f = g4mi.Fno(max_batch=2, seed=4, **sh)
    rng = np.random.default_rng(6)
    p = f.get_params().astype(np.float64)
    for name, o, n in f.param_blocks():
        if name.startswith("R"):
            p[o:o + n] = rng.normal(size=n) * 0.3
        elif name.startswith("b"):
            p[o:o + n] = rng.normal(size=n) * 0.1
    f.set_params(p.astype(np.float32))
    N = sh["ngrid"]
    inp = rng.normal(size=(2, N * N, 4)).astype(np.float32)
    tgt = rng.normal(size=(2, N * N, 2)).astype(np.float32)
    loss, g = f.loss_grad(inp, tgt)
    rl, rg, _ = R.loss_and_grad(p, inp, tgt, N, sh["width"], sh["modes"], sh["layers"],
                                sh["hidden"])
    _log(f"loss GPU {loss:.6e} ref {rl:.6e}")
    assert abs(loss / rl - 1) < 2e-2
    worst = 0.0
    for name, o, n in f.param_blocks():
        a, b = g[o:o + n].astype(np.float64), rg[o:o + n]
        e = _rel(a, b)
        cos = float(a @ b / max(np.linalg.norm(a) * np.linalg.norm(b), 1e-300))
        worst = max(worst, e)
        _log(f"  {name:4s} n={n:6d} rel L2 {e:.3e} cos {cos:.6f}")
        assert e < 5e-2 and cos > 0.998, name
    f.close()


def _capture_pairs(fno, nslices=16, npart=512, steps=60, stride=1, ngrid=33):
    """Mixed-form run of a small SASE window on ngrid^2; every step's ADI
    pairs of every stride-th slice go to the network's training set."""
    sim = H.make_state([npart] * nslices, ngrid=ngrid, gridmax=0.3e-3, seed_power=5e3,
                       longrange=True)
    ctx = H.to_gpu(sim, precision=g4mi.FP32)
    fno.reserve(fno.pair_count() + nslices * steps // stride + 1)
    ctx.attach_fno(fno, 1, every=stride)
    st = H.steps()
    ctx.run(H.gsteps(st[:steps]), cases.GAMMA0, record=())
    ctx.attach_fno(None, 0)
    return ctx, sim, st


def test_fno_training_on_adi_pairs_gains_over_baselines():
    """Pairs captured from 60 ADI steps of a 16-slice SASE window; training
    with Adam; the hold-out loss must fall below a fifth of the untrained
    network's and below a fifth of the field-unchanged prediction's (zero
    update, whose loss is the mean squared target)."""
    sh = dict(SMALL, width=16, modes=12)
    f = g4mi.Fno(max_batch=16, seed=7, **sh)
    ctx, _, _ = _capture_pairs(f)
    n = f.pair_count()
    nh = n // 8
    base = f.eval(n - nh, n, 16)
    tl, hl = f.train(epochs=20, batch=16, lr=3e-3, nholdout=nh, seed=2)
    # the field-unchanged prediction: a network with every parameter zero outputs 0
    p = f.get_params()
    f.set_params(np.zeros_like(p))
    zero = f.eval(n - nh, n, 16)
    f.set_params(p)
    _log(f"pairs {n} (hold-out {nh}); hold-out loss: untrained {base:.4e}, trained {hl[-1]:.4e}, "
         f"field-unchanged {zero:.4e}; train loss {tl[0]:.3e} -> {tl[-1]:.3e}")
    assert hl[-1] < 0.2 * base and hl[-1] < 0.2 * zero
    ctx.close()
    f.close()


def test_fno_replace_mode_falls_back_to_adi():
    """A trained network advances the field of a mixed-form run; the ADI
    runs on every 4th step and whenever the energy-balance residual exceeds
    the threshold.  Reported: the fallback rate, the mean residual of both,
    and P(z) against the FP64 oracle; asserted: the cadence and a finite
    run within 25 % of the oracle's P(z)."""
    sh = dict(SMALL, width=16, modes=12)
    f = g4mi.Fno(max_batch=16, seed=7, **sh)
    ctx, _, _ = _capture_pairs(f)
    f.train(epochs=20, batch=16, lr=3e-3, nholdout=0, seed=2)
    ctx.close()
    sim = H.make_state([512] * 16, ngrid=33, gridmax=0.3e-3, seed_power=5e3, longrange=True)
    ctx = H.to_gpu(sim, precision=g4mi.FP32)
    ctx.attach_fno(f, 2, every=4, threshold=1e-3)
    st = H.steps()
    n = 60
    g = ctx.run(H.gsteps(st[:n]), cases.GAMMA0, record=("power",))
    r = sim.run(st, cases.GAMMA0, record=("power",), nsteps=n)
    s = ctx.fno_stats()
    j = ctx.fno_judge()
    pg, pr = g["power"][1:].sum(axis=1), r["power"][1:].sum(axis=1)
    e = float(np.max(np.abs(pg / pr - 1)))
    _log(f"replace mode: {s}; judge steps {j}; fallback rate {s['adi_steps'] / n:.2f}; sum P(z) max "
         f"rel err vs FP64 oracle {e:.3e}")
    assert s["fno_steps"] + s["adi_steps"] == n
    assert s["adi_steps"] >= n // 4
    # every 4th step (0, 4, ..., 56) compares the network's step with the ADI's
    assert j["judge_steps"] == n // 4
    assert 0 < j["rel_field_err"] < 1 and np.isfinite(j["rel_update_err"])
    assert np.all(np.isfinite(pg)) and e < 0.25
    ctx.close()
    f.close()
