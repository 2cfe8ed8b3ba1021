"""GPU tests of the mixed-precision form (FP32 particles and field, FP64
slice phase base, FP64 residual correction in the ADI, FP64 reductions).

The mixed form is not bit-compared: each test states its error against the
FP64 oracle on the same inputs and bounds it.  The end-to-end bar is the
one the FP32 form is built to: P(z) and |b(z)| within 1 % of the FP64
oracle."""
import ctypes as C
import sys

import numpy as np
import pytest

import cases
import g4mi
import harness as H
import oracle_py as O

pytestmark = pytest.mark.gpu


def _log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def st():
    return H.steps()


def _field_err(ctx, sim):
    f, first = ctx.download_field()
    assert first == sim.s.first
    return float(np.max(np.abs(f - sim.field)) / np.max(np.abs(sim.field)))


@pytest.mark.parametrize("mode", [2, 1, 0])
@pytest.mark.parametrize("refine", [1, 0])
def test_mixed_adi_against_oracle(st, refine, mode):
    """Field-only steps (zero current: no source) over undulator, quadrupole
    and drift step sizes (band half widths 3, 4, 6 on 151^2): the FP32 band
    ADI, with and without the FP64 residual correction, in each kernel
    (2: whole-width streaming, 1: streaming strips, 0: 32x32 tiles), against
    the FP64 Thomas ADI of the oracle.  Bound: the FP32 rounding of the
    stored field per step (2^-24 relative) times the step count, with
    margin."""
    _log(f"mixed ADI vs oracle, refine={refine}, kernel {mode}: 4 slices x 256, 151^2, steps 120..130")
    # steady state: no slippage, so the seed stays on the grid for every step
    sim = H.make_state([256] * 4, seed_power=5e3, timerun=False)
    sim.current[:] = 0.0
    ctx = H.to_gpu(sim, precision=g4mi.FP32)
    ctx.set_schedule("adi_refine", refine)
    ctx.set_schedule("mixed_adi", mode)
    idx = list(range(120, 131))
    errs = []
    for i in idx:
        O.lib().orc_step_all(C.byref(sim.s), C.byref(st[i]), cases.GAMMA0)
        ctx.step(H.gsteps(st)[i], cases.GAMMA0)
        errs.append(_field_err(ctx, sim))
    sched = ctx.schedule()
    assert sched["adi"] == {2: "band32-wide", 1: "band32-stream", 0: "band32-tile"}[mode]
    _log(f"  dz {[round(st[i].dz, 3) for i in idx]}\n  field rel err per step {['%.2e' % e for e in errs]}"
         f"\n  schedule {sched}")
    bound = (2.0 if refine else 8.0) * 2.0 ** -24 * len(idx)
    assert np.all(np.isfinite(errs)), errs
    assert max(errs) < bound, (max(errs), bound)
    ctx.close()


def test_mixed_adi_kernels_agree_with_source(st):
    """The three ADI kernels of the mixed form on a SASE window with a
    source, long-range space charge and slippage (10 slices x 1024, 131^2:
    the whole-width kernel then has idle lanes and a partly filled last
    wavefront) over 40 steps: records and final field of the whole-width
    kernel and of the strips against the tile kernel.  All three round the
    same FP32 band products in a different order only, so they agree to
    a few FP32 ulps per step."""
    _log("mixed ADI kernels agree: 10 slices x 1024, 131^2, 40 steps, modes 2, 1, 0")
    out = {}
    for mode in (2, 1, 0):
        sim = H.make_state([1024] * 10, ngrid=131, gridmax=0.4e-3, longrange=True, seed_power=1e4)
        ctx = H.to_gpu(sim, precision=g4mi.FP32)
        ctx.set_schedule("mixed_adi", mode)
        r = ctx.run(H.gsteps(st[120:160]), cases.GAMMA0, record=("power", "bunching"))
        f, first = ctx.download_field()
        out[mode] = (r, f, first)
        ctx.close()
    r0, f0, q0 = out[0]
    for mode in (2, 1):
        r, f, q = out[mode]
        assert q == q0
        ep = float(np.max(np.abs(r["power"] - r0["power"])) / np.max(np.abs(r0["power"])))
        eb = float(np.max(np.abs(r["bunching"] - r0["bunching"])) / np.max(np.abs(r0["bunching"])))
        ef = float(np.max(np.abs(f - f0)) / np.max(np.abs(f0)))
        _log(f"  kernel {mode} vs tiles: power {ep:.3e}, bunching {eb:.3e}, field {ef:.3e}")
        assert ep < 1e-4 and eb < 1e-4 and ef < 1e-4


def test_mixed_free_push_offset_form(st):
    """Zero field and current: the FP32 offset push (gamma = gref + dg,
    theta = base + dth, phase rate xks g'(eps) (eps_ref - eps)) against the
    FP64 oracle over the first module (125 steps) plus the drift after it.
    The stored offsets take one stochastically rounded update per step, so
    the theta error is a random walk of ~ulp(2 pi)/2 = 2.4e-7 rad per step:
    std ~ 2.4e-7 sqrt(131) ~ 3e-6, the largest of 4096 particles ~5 std;
    bounds 5e-5 on the largest and 1.5e-5 on the mean.  Gamma: constant, to
    the FP32 rounding of dg."""
    _log("mixed free push vs oracle: 8 slices x 512, 131 steps")
    sim = H.make_state([512] * 8)
    sim.current[:] = 0.0
    ctx = H.to_gpu(sim, precision=g4mi.FP32)
    n = 131
    ctx.run(H.gsteps(st[:n]), cases.GAMMA0, record=())
    for i in range(n):
        O.lib().orc_step_all(C.byref(sim.s), C.byref(st[i]), cases.GAMMA0)
    b = ctx.download_beam()
    dth = float(np.max(np.abs(b["theta"] - sim.theta)))
    dg = float(np.max(np.abs(b["gamma"] - sim.gamma)))
    dx = float(np.max(np.abs(b["x"] - sim.x)) / np.max(np.abs(sim.x)))
    mth = float(np.mean(b["theta"] - sim.theta))
    _log(f"  max |dtheta| {dth:.3e} rad (mean {mth:.3e}), max |dgamma| {dg:.3e}, x rel {dx:.3e}")
    assert dth < 5e-5 and abs(mth) < 1.5e-5
    assert dg < 1e-3          # 2 ulp of FP32 at |dg| ~ 5 (gamma spread 1e-4)
    # x: 262 FP32 linear-map half steps, each rounding at 2^-24 relative
    assert dx < 3e-5
    ctx.close()


def _mixed_vs_oracle(sim, st, n, record=("power", "bunching")):
    ctx = H.to_gpu(sim, precision=g4mi.FP32)
    g = ctx.run(H.gsteps(st[:n]), cases.GAMMA0, record=record)
    r = sim.run(st, cases.GAMMA0, record=record, nsteps=n)
    return ctx, g, r


def test_mixed_steady_state_seeded_within_one_percent(st):
    """Config 0 (one slice of 8192, 5 kW seed, 151^2) over all 509 steps in
    the mixed form: P(z) and |b(z)| within 1 % of the FP64 oracle."""
    _log("mixed config 0, 509 steps, vs FP64 oracle")
    sim = H.make_state([8192], seed_power=5e3, timerun=False)
    ctx, g, r = _mixed_vs_oracle(sim, st, len(st))
    ep = np.abs(g["power"][:, 0] / r["power"][:, 0] - 1)
    eb = np.abs(g["bunching"][:, 0] / r["bunching"][:, 0] - 1)
    mask = r["bunching"][:, 0] > 1e-3  # |b| of the quiet start is ~1e-17 (noise floor)
    _log(f"  P(z) max rel err {ep.max():.3e} (end {ep[-1]:.3e}, P_end {r['power'][-1, 0]:.4e}); "
         f"|b(z)| max rel err {eb[mask].max():.3e} over {mask.sum()} steps with |b| > 1e-3")
    assert ep.max() < 1e-2
    assert eb[mask].max() < 1e-2
    ctx.close()


def test_mixed_sase_small_within_one_percent(st):
    """Time-dependent SASE on a reduced window (24 slices x 1024, shot noise
    seed 7, long-range space charge, slippage, 63^2): the slice-summed P(z)
    and the mean |b(z)| within 1 % of the FP64 oracle over all 509 steps,
    and the final field within 1 % (relative to its maximum)."""
    _log("mixed time-dependent SASE, 24 slices, 509 steps, vs FP64 oracle")
    sim = H.make_state([1024] * 24, ngrid=63, gridmax=0.3e-3, longrange=True)
    ctx, g, r = _mixed_vs_oracle(sim, st, len(st), record=("power", "bunching"))
    pg, pr = g["power"][1:].sum(axis=1), r["power"][1:].sum(axis=1)
    bg, br = g["bunching"].mean(axis=1), r["bunching"].mean(axis=1)
    ep = float(np.max(np.abs(pg / pr - 1)))
    eb = float(np.max(np.abs(bg / br - 1)))
    f, first = ctx.download_field()
    ef = float(np.max(np.abs(f - sim.field)) / np.max(np.abs(sim.field)))
    _log(f"  sum P(z) max rel err {ep:.3e}  mean |b(z)| max rel err {eb:.3e}  final field {ef:.3e}"
         f"  (P_end {pr[-1]:.4e})")
    assert first == sim.s.first
    assert ep < 1e-2 and eb < 1e-2 and ef < 1e-2
    ctx.close()


def test_mixed_graph_replay_matches_direct_launch(st):
    """Every replay of the captured step graphs (undulator and drift step
    signatures) against direct launches from the same initial state.  The
    deposit and the space-charge histograms add with FP32 LDS atomics, whose
    order varies run to run, so the bound is FP32 rounding, not equality."""
    _log("mixed graph replay vs direct launch, 140 steps, 6 slices x 512")
    recs = []
    for graphs in (1, 0):
        sim = H.make_state([512] * 6, ngrid=63, gridmax=0.3e-3, longrange=True,
                           sc=(1, 0, 100, 0.0))
        ctx = H.to_gpu(sim, precision=g4mi.FP32)
        ctx.set_schedule("graphs", graphs)
        out = {"power": [], "bunching": []}
        for k in range(0, 140, 35):  # eager work and synchronisation between runs
            r = ctx.run(H.gsteps(st[k:k + 35]), cases.GAMMA0, record=("power", "bunching"))
            ctx.sync()
            out["power"].append(r["power"])
            out["bunching"].append(r["bunching"])
        sched = ctx.schedule()
        recs.append({k: np.concatenate(v) for k, v in out.items()})
        ctx.close()
        _log(f"  graphs={graphs}: {sched}")
    for k in ("power", "bunching"):
        a, b = recs[0][k], recs[1][k]
        e = float(np.max(np.abs(a - b)) / np.max(np.abs(b)))
        _log(f"  {k}: max rel diff graph vs direct {e:.3e}")
        assert e < 1e-4


def test_mixed_persistent_step_matches_kernels(st):
    """The whole step as one persistent cooperative launch (grid barriers
    between front, scan, main, ADI, slippage, records) against the same
    steps as separate kernels from the same state: 60 steps of a 40-slice
    window with long- and short-range space charge and slippage.  Records
    and final field within FP32 rounding (the deposit and histograms add
    with LDS atomics whose order varies)."""
    _log("mixed persistent step vs kernels, 60 steps, 40 slices x 512, 63^2")
    out = []
    for pers in (1, 0):
        sim = H.make_state([512] * 40, ngrid=63, gridmax=0.3e-3, longrange=True, sc=(1, 0, 100, 0.0))
        ctx = H.to_gpu(sim, precision=g4mi.FP32)
        ctx.set_schedule("persistent", pers)
        r = ctx.run(H.gsteps(st[:60]), cases.GAMMA0, record=("power", "bunching"))
        f, first = ctx.download_field()
        sched = ctx.schedule()
        assert sched["persistent"] == pers, sched
        out.append((r, f, first))
        ctx.close()
    (ra, fa, qa), (rb, fb, qb) = out
    assert qa == qb
    for k in ("power", "bunching"):
        e = float(np.max(np.abs(ra[k] - rb[k])) / np.max(np.abs(rb[k])))
        _log(f"  {k}: max rel diff persistent vs kernels {e:.3e}")
        assert e < 1e-4
    ef = float(np.max(np.abs(fa - fb)) / np.max(np.abs(fb)))
    _log(f"  final field: max diff / max {ef:.3e}")
    assert ef < 1e-4


def test_mixed_fused_front_matches_separate_kernels(st):
    """The main kernel running the next step's front for its slice (the
    default) against every front as its own kernel: 60 steps of a 40-slice
    window with long- and short-range space charge and slippage, as runs of
    1, 20 and 39 steps (a run's last step leaves the particles in the
    reference's order, the next run starts with its own front), with 256
    threads per slice (the front's partition) and with 384.  Records, final
    field and the downloaded beam within FP32 rounding (the deposit and
    histograms add with LDS atomics whose order varies)."""
    for nt in (256, 384):
        _log(f"mixed fused front vs separate kernels, {nt} threads, 60 steps, 40 slices x 512, 63^2")
        out = []
        for fuse in (1, 0):
            sim = H.make_state([512] * 40, ngrid=63, gridmax=0.3e-3, longrange=True, sc=(1, 0, 100, 0.0))
            ctx = H.to_gpu(sim, precision=g4mi.FP32)
            ctx.set_schedule("mixed_fuse_front", fuse)
            ctx.set_schedule("main_nt", nt)
            rec = {"power": [], "bunching": []}
            for a, b in ((0, 1), (1, 21), (21, 60)):
                r = ctx.run(H.gsteps(st[a:b]), cases.GAMMA0, record=("power", "bunching"))
                for k in rec:
                    rec[k].append(r[k])
            f, first = ctx.download_field()
            beam = ctx.download_beam()
            sched = ctx.schedule()
            assert sched["fuse_front"] == fuse, sched
            out.append(({k: np.concatenate(v) for k, v in rec.items()}, f, first, beam))
            ctx.close()
        (ra, fa, qa, ba), (rb, fb, qb, bb) = out
        assert qa == qb
        for k in ("power", "bunching"):
            e = float(np.max(np.abs(ra[k] - rb[k])) / np.max(np.abs(rb[k])))
            _log(f"  {k}: max rel diff fused vs separate {e:.3e}")
            assert e < 1e-4
        ef = float(np.max(np.abs(fa - fb)) / np.max(np.abs(fb)))
        _log(f"  final field: max diff / max {ef:.3e}")
        assert ef < 1e-4
        for k in ("x", "y", "px", "py", "gamma", "theta"):
            sc = max(float(np.max(np.abs(bb[k] - (np.mean(bb[k]) if k == "gamma" else 0.0)))), 1e-300)
            e = float(np.max(np.abs(ba[k] - bb[k]))) / sc
            _log(f"  beam {k}: max diff / scale {e:.3e}")
            assert e < 1e-4, k


def test_mixed_fused_front_with_tile_overflow(st):
    """A beam with a transverse momentum offset (px + 50, about 3.4 cells
    per half step at 63^2): the second half step carries particles out of
    the deposit tile (the first half step's footprint padded by 2 cells), so
    every slice takes the two-pass deposit, which with the next front in the
    main loop reads the post-t2 positions kept aside.  Three undulator
    steps as one run, next front in the loop against separate front kernels:
    records, field and beam within FP32 rounding."""
    _log("mixed fused front, deposit tile overflow: 6 slices x 512, 63^2, 3 steps")
    out = []
    for fuse in (1, 0):
        sim = H.make_state([512] * 6, ngrid=63, gridmax=0.3e-3, longrange=True, sc=(1, 0, 100, 0.0),
                           seed_power=1e4)
        ctx = H.to_gpu(sim, precision=g4mi.FP32)
        beam = {k: np.array(getattr(sim, k)) for k in ("gamma", "theta", "x", "y", "px", "py")}
        beam["px"] += 50.0
        ctx.upload_beam(beam, sim.current)
        ctx.set_schedule("mixed_fuse_front", fuse)
        ctx.set_schedule("main_nt", 256)
        ctx.set_schedule("main_lds", 1)
        r = ctx.run(H.gsteps(st[:3]), cases.GAMMA0, record=("power", "bunching"))
        f, first = ctx.download_field()
        out.append((r, f, first, ctx.download_beam()))
        ctx.close()
    (ra, fa, qa, ba), (rb, fb, qb, bb) = out
    assert qa == qb
    x0 = np.array(bb["x"])
    moved = float(np.mean(x0)) / (2 * 0.3e-3 / 62)
    _log(f"  beam centroid moved {moved:.1f} cells in 3 steps")
    assert moved > 15
    for k in ("power", "bunching"):
        e = float(np.max(np.abs(ra[k] - rb[k])) / np.max(np.abs(rb[k])))
        _log(f"  {k}: max rel diff fused vs separate {e:.3e}")
        assert e < 1e-4
    ef = float(np.max(np.abs(fa - fb)) / np.max(np.abs(fb)))
    _log(f"  final field: max diff / max {ef:.3e}")
    assert ef < 1e-4
    for k in ("x", "y", "px", "py", "gamma", "theta"):
        sc = max(float(np.max(np.abs(bb[k] - (np.mean(bb[k]) if k == "gamma" else 0.0)))), 1e-300)
        e = float(np.max(np.abs(ba[k] - bb[k]))) / sc
        assert e < 1e-4, (k, e)


def test_mixed_autotuner_picks_a_schedule(st):
    """The first run of 18 or more steps times every main-kernel shape (4, 6,
    8 wavefronts per slice x 68 or 44 KiB of LDS; 4, 6 x 32 KiB), the three ADI kernels, the
    graph of kernels and the persistent launch, and keeps the fastest; the
    run's records stay within 1 % of the FP64 oracle."""
    _log("mixed autotuner: 24 steps, 12 slices x 1024, 63^2")
    sim = H.make_state([1024] * 12, ngrid=63, gridmax=0.3e-3, longrange=True, sc=(1, 0, 100, 0.0))
    ctx = H.to_gpu(sim, precision=g4mi.FP32)
    n = 24
    g = ctx.run(H.gsteps(st[:n]), cases.GAMMA0, record=("power", "bunching"))
    sched = ctx.schedule()
    _log(f"  schedule {sched}")
    assert sched["tuned"] == 1
    tun = sched["tuning"]
    assert len(tun["main_ms"]) == 8 and all(v > 0 for v in tun["main_ms"].values())
    assert tun["adi_ms"]["wide"] > 0 and tun["adi_ms"]["strips"] > 0 and tun["step_ms"]["graph"] > 0
    assert sched["adi"] in ("band32-wide", "band32-stream", "band32-tile")
    assert sched["main_nt"] in (256, 384, 512) and sched["main_lds_kib"] in (32, 44, 68)
    r = sim.run(st, cases.GAMMA0, record=("power", "bunching"), nsteps=n)
    pg, pr = g["power"][1:].sum(axis=1), r["power"][1:].sum(axis=1)
    e = float(np.max(np.abs(pg / pr - 1)))
    _log(f"  sum P(z) max rel err vs oracle {e:.3e}")
    assert e < 1e-2
    ctx.close()


def test_mixed_main_kernel_shapes_agree(st):
    """Every main-kernel shape the tuner chooses from (4, 6, 8 wavefronts per
    slice x 68 or 44 KiB of LDS; 4, 6 x 32 KiB) on 6 slices x 4096 at 255^2
    with space charge (one mode of 100 radial cells), 30 steps: the slices'
    padded deposit footprints (~47 x 48 cells) fit the 44 and 68 KiB tiles
    (2766 and 4302 cells) but not the 32 KiB one (1998 cells), whose shapes
    take the banded two-pass deposit.  Records and field agree to FP32
    rounding (the LDS atomics add in varying order)."""
    _log("mixed main-kernel shapes: 6 slices x 4096, 255^2, 30 steps, 8 shapes")
    ref = None
    for nt, lds in ((384, 0), (256, 0), (512, 0), (384, 1), (256, 1), (512, 1), (256, 2), (384, 2)):
        sim = H.make_state([4096] * 6, ngrid=255, gridmax=0.5e-3, longrange=True, sc=(1, 0, 100, 0.0),
                           seed_power=1e4)
        ctx = H.to_gpu(sim, precision=g4mi.FP32)
        ctx.set_schedule("main_nt", nt)
        ctx.set_schedule("main_lds", lds)
        r = ctx.run(H.gsteps(st[:30]), cases.GAMMA0, record=("power", "bunching"))
        f, _ = ctx.download_field()
        sched = ctx.schedule()
        ctx.close()
        assert sched["main_nt"] == nt and sched["main_lds_kib"] == (68, 44, 32)[lds]
        if ref is None:
            ref = (r, f)
            continue
        ep = float(np.max(np.abs(r["power"] - ref[0]["power"])) / np.max(np.abs(ref[0]["power"])))
        eb = float(np.max(np.abs(r["bunching"] - ref[0]["bunching"])) / np.max(np.abs(ref[0]["bunching"])))
        ef = float(np.max(np.abs(f - ref[1])) / np.max(np.abs(ref[1])))
        _log(f"  {nt} threads, {sched['main_lds_kib']} KiB: power {ep:.2e}, bunching {eb:.2e}, field {ef:.2e}")
        assert ep < 1e-4 and eb < 1e-4 and ef < 1e-4


def test_mixed_upload_download_and_sort_index():
    """Offset form round trip: gamma and theta come back to FP32 rounding of
    their offsets; the sort target slice equals floor(theta / slen) of the
    downloaded theta."""
    sim = H.make_state([300, 16, 516, 64])
    ctx = H.to_gpu(sim, precision=g4mi.FP32)
    b = ctx.download_beam()
    assert np.max(np.abs(b["gamma"] - sim.gamma)) <= 1e-3
    assert np.max(np.abs(b["theta"] - sim.theta)) <= 1e-6
    for k in ("x", "y", "px", "py"):
        ref = getattr(sim, k)
        assert np.max(np.abs(b[k] - ref) / (np.abs(ref) + 1e-300)) <= 6e-8
    slen = 2 * np.pi
    atar, moved = ctx.sort_index(slen)
    want = np.floor(b["theta"] * (1.0 / slen)).astype(int)
    assert np.array_equal(atar, want)
    assert moved == int(np.count_nonzero(want))
    ctx.close()


def test_mixed_diag_aux_and_global(st):
    """The auxiliary and global diagnostics of the mixed form against numpy
    on its own downloaded state (FP64 reductions of the same values), and
    its space-charge fields against the FP64 oracle's after the same steps
    (FP32 moments and solve: 1e-4 and 1e-3)."""
    sim = H.make_state([1024, 512, 768, 256], seed_power=5e3, longrange=True, sc=(1, 0, 100, 0.0),
                       currents=[3000.0, 2000.0, 2500.0, 1000.0])
    ctx = H.to_gpu(sim, precision=g4mi.FP32)
    ctx.run(H.gsteps(st[:3]), cases.GAMMA0, record=())
    for i in range(3):
        O.lib().orc_step_all(C.byref(sim.s), C.byref(st[i]), cases.GAMMA0)
    d = ctx.diag()
    b = ctx.download_beam()
    f, first = ctx.download_field()
    ref = H.aux_reference(b, f, first, sim.off, sim.current, sim.s.ngrid, sim.s.dgrid, cases.LAMBDA_R)
    H.check_aux(d, ref, b, 1e-6)
    assert H.rel_err(d["lscfield"], sim.longESC, 1e-12) < 1e-4
    assert H.rel_err(d["sscfield"], sim.sc_efield, 1e-12) < 1e-3
    ctx.close()


def test_mixed_rejects_split_calls(st):
    sim = H.make_state([64] * 2)
    ctx = H.to_gpu(sim, precision=g4mi.FP32)
    with pytest.raises(g4mi.G4miError):
        ctx.beam_track(H.gsteps(st)[0], cases.GAMMA0, st[0].dz)
    ctx.close()
