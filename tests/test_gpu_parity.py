"""GPU parity: the HIP FP64 plain form against the FP64 CPU oracle on the
same inputs, through the C-ABI.  Tolerances: index work (slippage ring
offset, sort index) bit-exact; FP64 floating point within 1e-10 relative
(P(z), |b(z)|, final field), with a per-kernel check tighter than that.
The only sources of difference are the device's sin/cos/log/exp (an ulp
from glibc's) and reduction order in sums."""
import ctypes as C
import sys

import numpy as np
import pytest

import cases
import g4mi
import harness as H
import oracle_py as O

pytestmark = pytest.mark.gpu

BEAM = ("gamma", "theta", "x", "y", "px", "py")


def _log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _beam_close(ctx, sim, rtol):
    b = ctx.download_beam()
    worst = {}
    for k in BEAM:
        ref = getattr(sim, k)
        scale = np.maximum(np.abs(ref), 1e-3 * np.max(np.abs(ref)) + 1e-300)
        worst[k] = float(np.max(np.abs(b[k] - ref) / scale))
    assert max(worst.values()) <= rtol, worst
    return worst


@pytest.fixture(scope="module")
def st():
    return H.steps()


def test_transverse_parity(st):
    """Beam track at zero field and current: the transverse half steps
    (drift, quadrupole, undulator focusing) and the free push."""
    _log("transverse parity: 8 slices x 512, undulator/drift/quad/drift steps")
    sim = H.make_state([512] * 8)
    sim.current[:] = 0.0
    ctx = H.to_gpu(sim)
    gst = H.gsteps(st)
    for idx in (0, 125, 126, 127):
        O.lib().orc_beam_track(C.byref(sim.s), C.byref(st[idx]), cases.GAMMA0, st[idx].dz)
        ctx.beam_track(gst[idx], cases.GAMMA0, st[idx].dz)
    w = _beam_close(ctx, sim, 1e-14)
    _log(f"  worst rel diff {w}")
    ctx.close()


def test_push_parity_with_field(st):
    _log("push parity: 8 slices x 1024 with a 50 MW seed field, 5 undulator steps")
    sim = H.make_state([1024] * 8, seed_power=5e7)
    ctx = H.to_gpu(sim)
    gst = H.gsteps(st)
    for i in range(5):
        O.lib().orc_beam_track(C.byref(sim.s), C.byref(st[i]), cases.GAMMA0, st[i].dz)
        ctx.beam_track(gst[i], cases.GAMMA0, st[i].dz)
    w = _beam_close(ctx, sim, 1e-13)
    _log(f"  worst rel diff {w}")
    ctx.close()


def test_field_track_parity(st):
    _log("field parity: deposit + ADI, 8 slices x 2048, 151^2, 3 steps")
    sim = H.make_state([2048] * 8, seed_power=5e3)
    ctx = H.to_gpu(sim)
    gst = H.gsteps(st)
    for i in range(3):
        O.lib().orc_field_advance(C.byref(sim.s), C.byref(st[i]), st[i].dz)
        ctx.field_track(gst[i], st[i].dz)
    f, first = ctx.download_field()
    assert first == sim.s.first
    err = np.max(np.abs(f - sim.field)) / np.max(np.abs(sim.field))
    _log(f"  field max diff / max |E| = {err:.3e}")
    assert err < 1e-13


def test_fused_step_matches_separate_kernels(st):
    """The FP64 step's particle passes as two fused kernels (front; space
    charge + push + R56 + transverse + beam sums + deposit) against the
    separate kernels from the same state: 60 steps (undulator, drift and
    quadrupole) of an uneven 10-slice SASE window with long- and short-range
    space charge.  Both are the reference's arithmetic; they differ only in
    the LDS-atomic order of the histograms and the deposit, so records,
    particles and field agree to 1e-12, and each matches the oracle."""
    _log("fused vs separate FP64 particle kernels: 10 uneven slices, LSC, 60 steps")
    counts = [512, 1024, 256, 2048, 768, 1024, 512, 256, 1536, 1024]
    out = []
    for fused in (1, 0):
        sim = H.make_state(counts, ngrid=63, gridmax=0.3e-3, longrange=True, sc=(1, 0, 100, 0.0),
                           currents=[1000.0 + 200 * k for k in range(len(counts))])
        ctx = H.to_gpu(sim)
        ctx.set_schedule("fp64_fused", fused)
        idx = list(range(120, 180))
        r = ctx.run(H.gsteps([st[i] for i in idx]), cases.GAMMA0, record=("power", "bunching"))
        sched = ctx.schedule()
        assert sched["fp64_fused"] == fused
        b = ctx.download_beam()
        f, first = ctx.download_field()
        out.append((r, b, f, first))
        ctx.close()
    rs = sim.run([st[i] for i in idx], cases.GAMMA0, record=("power", "bunching"))
    (ra, ba, fa, qa), (rb, bb, fb, qb) = out
    assert qa == qb == sim.s.first
    for k in ("power", "bunching"):
        e = H.rel_err(ra[k], rb[k])
        eo = H.rel_err(ra[k], rs[k])
        _log(f"  {k}: fused vs separate {e:.3e}, fused vs oracle {eo:.3e}")
        assert e < 1e-12 and eo < 1e-10
    for k in BEAM:
        e = float(np.max(np.abs(ba[k] - bb[k])) / (np.max(np.abs(bb[k])) + 1e-300))
        assert e < 1e-12, (k, e)
    ef = float(np.max(np.abs(fa - fb)) / np.max(np.abs(fb)))
    _log(f"  final field fused vs separate {ef:.3e}")
    assert ef < 1e-12


@pytest.mark.parametrize("fused", [1, 0])
def test_fp64_runs_are_bit_reproducible(st, fused):
    """The space-charge histograms and the deposit sum in LDS in 64-bit
    fixed point (integer atomics commute), so two runs of the FP64 form from
    the same state give the same bits: records, particles and field, over
    40 steps with long- and short-range space charge."""
    _log(f"FP64 bit reproducibility, fused={fused}: 6 uneven slices, LSC, 40 steps")
    counts = [2048, 512, 4096, 1024, 256, 3072]
    out = []
    for _ in range(2):
        sim = H.make_state(counts, ngrid=63, gridmax=0.3e-3, longrange=True, sc=(1, 0, 100, 0.0),
                           currents=[1500.0 + 300 * k for k in range(len(counts))], seed_power=1e4)
        ctx = H.to_gpu(sim)
        ctx.set_schedule("fp64_fused", fused)
        r = ctx.run(H.gsteps([st[i] for i in range(120, 160)]), cases.GAMMA0, record=("power", "bunching"))
        b = ctx.download_beam()
        f, first = ctx.download_field()
        out.append((r, b, f, first))
        ctx.close()
    (ra, ba, fa, qa), (rb, bb, fb, qb) = out
    assert qa == qb
    for k in ("power", "bunching"):
        assert np.array_equal(ra[k], rb[k]), k
    for k in BEAM:
        assert np.array_equal(ba[k], bb[k]), k
    assert np.array_equal(fa, fb)


def test_deposit_corner_cell_and_mixed_groups(st):
    """The deposit combines the four lanes of a group when they share a
    cell.  Groups here: all four in the corner cell (0, 0); one in the
    corner and three elsewhere; two and two; four different cells; four in
    one interior cell; and a tail group cut short by the slice's end."""
    _log("deposit parity: corner cell (0,0) and mixed lane groups, 3 slices")
    sim = H.make_state([256, 254, 64], seed_power=5e3)
    # the loader fills whole beamlets of nbins = 4: the last two slots of
    # the 254-particle slice take the coordinates of the two before them
    e1 = int(sim.off[2])
    for k in ("gamma", "theta", "px", "py"):
        getattr(sim, k)[e1 - 2:e1] = getattr(sim, k)[e1 - 4:e1 - 2]
    gm, dg = sim.s.gridmax, sim.s.dgrid
    corner = -gm + 0.3 * dg                    # inside cell 0
    rng = np.random.default_rng(21)
    n = int(sim.off[-1])
    x = rng.normal(size=n) * 2e-5
    y = rng.normal(size=n) * 2e-5
    cells = [(corner, corner)] * 4, [(corner, corner), (0.0, 0.0), (1e-5, 0.0), (0.0, 2e-5)], \
        [(corner, corner), (corner, corner), (3e-5, 3e-5), (3e-5, 3e-5)], \
        [(0.0, 0.0), (dg, 0.0), (0.0, dg), (dg, dg)], [(2.1e-6, -1.3e-6)] * 4
    for a in range(sim.ns):
        base = int(sim.off[a])
        for gi, grp in enumerate(cells):
            for k, (px_, py_) in enumerate(grp):
                x[base + 4 * gi + k] = px_ + 1e-9 * k
                y[base + 4 * gi + k] = py_ + 1e-9 * k
    sim.x[:], sim.y[:] = x, y
    ctx = H.to_gpu(sim)
    gst = H.gsteps(st)
    for i in range(2):
        O.lib().orc_field_advance(C.byref(sim.s), C.byref(st[i]), st[i].dz)
        ctx.field_track(gst[i], st[i].dz)
    f, first = ctx.download_field()
    err = np.max(np.abs(f - sim.field)) / np.max(np.abs(sim.field))
    ng = sim.s.ngrid
    cidx = [iy * ng + ix for iy in range(3) for ix in range(3)]    # the corner's 3x3 cells
    cref = sim.field[:, cidx]
    assert np.max(np.abs(cref)) > 0
    corner_err = np.max(np.abs(f[:, cidx] - cref)) / np.max(np.abs(cref))
    _log(f"  field max diff / max |E| = {err:.3e}, corner cell {corner_err:.3e}")
    assert err < 1e-13 and corner_err < 1e-12
    ctx.close()


@pytest.mark.parametrize("kernel", [0, 1, 2, 3, 4])
@pytest.mark.parametrize("delz", [0.015, 0.15, 5.0])
def test_adi_kernels_and_regimes(delz, kernel):
    """rtmp = dz/(4 ks dx^2) = 3e-3, 3e-2 and 1 on the 151^2 grid: the fused
    recurrence kernels use an 8-cell window, a 16-cell window, or are not
    eligible; the banded-inverse kernels a half width 7 (tiles and the
    whole-width stream), 12 (tiles only), or are not eligible.  Every
    eligible kernel agrees with the oracle's Thomas; an ineligible request
    falls back to the autotuner's choice."""
    _log(f"ADI regime delz={delz} kernel={kernel}")
    sim = H.make_state([1024] * 4, seed_power=5e7)
    ctx = H.to_gpu(sim)
    ctx.set_schedule("adi_kernel", kernel)
    st = O.Step(aw=cases.AW, ku=2 * np.pi / cases.LAMBDAU, ky=(2 * np.pi / cases.LAMBDAU) ** 2,
                dz=delz)
    for _ in range(3):
        O.lib().orc_field_advance(C.byref(sim.s), C.byref(st), delz)
        ctx.field_track(g4mi.steps_from([st])[0], delz)
    f, _ = ctx.download_field()
    err = np.max(np.abs(f - sim.field)) / np.max(np.abs(sim.field))
    sch = ctx.schedule()
    _log(f"  field max diff / max |E| = {err:.3e}  schedule {sch}")
    assert err < 1e-13
    assert sch["adi_window"] == {0.015: 8, 0.15: 16, 5.0: 0}[delz]
    assert sch["adi_band_m"] == {0.015: 7, 0.15: 12, 5.0: 0}[delz]
    if kernel == 4 and delz == 0.015:
        assert sch["adi"] == "band-wide"


@pytest.mark.parametrize("delz,m", [(0.35, 16), (0.8, 0)])
def test_adi_band_tiles_wide_halfwidths(delz, m):
    """Longer steps: rtmp = 0.07 on 151^2 needs a half width of 16 for the
    banded inverse to drop only entries below 2^-60 of the diagonal, the
    widest the 32x32-tile kernel takes (the parallel form for a single
    steady-state slice, whose Thomas sweeps run 151 lines deep); rtmp = 0.16
    would need 24, so the band is not eligible there and the request falls
    back to the tuner's choice.  Both against the oracle's Thomas over 3
    steps."""
    _log(f"banded ADI tiles, half width {m}: delz={delz}, 151^2, 4 slices")
    sim = H.make_state([1024] * 4, seed_power=5e7)
    ctx = H.to_gpu(sim)
    ctx.set_schedule("adi_kernel", 3)
    st = O.Step(aw=cases.AW, ku=2 * np.pi / cases.LAMBDAU, ky=(2 * np.pi / cases.LAMBDAU) ** 2, dz=delz)
    for _ in range(3):
        O.lib().orc_field_advance(C.byref(sim.s), C.byref(st), delz)
        ctx.field_track(g4mi.steps_from([st])[0], delz)
    f, _ = ctx.download_field()
    err = np.max(np.abs(f - sim.field)) / np.max(np.abs(sim.field))
    sch = ctx.schedule()
    _log(f"  field max diff / max |E| = {err:.3e}  schedule {sch}")
    assert sch["adi_band_m"] == m and (sch["adi"] == "band") == (m > 0)
    assert err < 1e-13
    ctx.close()


@pytest.mark.parametrize("ngrid", [151, 255, 63])
def test_adi_wide_stream_matches_band_tiles(ngrid):
    """The whole-width banded ADI evaluates the products of the 32x32-tile
    banded kernel; a column or row takes the symmetric form of the central
    band where its own wavefront (stream) or its whole tile (tiles) lies in
    [klo, khi], so the two differ by the rounding of that choice only: over
    6 undulator steps with a source the fields agree to 1e-14 and both
    match the oracle's Thomas."""
    _log(f"banded ADI: whole-width stream vs tiles, {ngrid}^2, 6 slices, 6 steps")
    st = H.steps()
    out = []
    for kernel in (3, 4):
        sim = H.make_state([512, 256, 1024, 768, 512, 256], ngrid=ngrid, gridmax=0.3e-3 * ngrid / 151,
                           seed_power=2e4)
        ctx = H.to_gpu(sim)
        ctx.set_schedule("adi_kernel", kernel)
        r = ctx.run(H.gsteps(st[118:124]), cases.GAMMA0, record=("power",))
        sch = ctx.schedule()
        assert sch["adi"] == ("band" if kernel == 3 else "band-wide"), sch
        f, first = ctx.download_field()
        out.append((r, f, first))
        ctx.close()
    rs = sim.run(st[118:124], cases.GAMMA0, record=("power",))
    (ra, fa, qa), (rb, fb, qb) = out
    assert qa == qb == sim.s.first
    ef = float(np.max(np.abs(fa - fb)) / np.max(np.abs(fa)))
    eo = float(np.max(np.abs(fb - sim.field)) / np.max(np.abs(fa)))
    _log(f"  field stream vs tiles {ef:.3e}, stream vs oracle {eo:.3e}")
    assert ef < 1e-14 and eo < 1e-12
    ep = H.rel_err(rb["power"], ra["power"])
    eo = H.rel_err(rb["power"], rs["power"])
    _log(f"  power stream vs tiles {ep:.3e}, stream vs oracle {eo:.3e}")
    assert ep < 1e-13 and eo < 1e-10


def test_adi_autotuner_picks_a_measured_variant():
    sim = H.make_state([512] * 16, seed_power=5e3)
    ctx = H.to_gpu(sim)
    st = H.steps()
    ctx.field_track(H.gsteps(st)[0], st[0].dz)
    sch = ctx.schedule()
    _log(f"  autotuned schedule {sch}")
    ms = sch["adi_tuned_ms"]
    assert sch["adi_kernel"] in (0, 1, 2, 3, 4) and ms[sch["adi_kernel"]] == min(m for m in ms if m > 0)
    O.lib().orc_field_advance(C.byref(sim.s), C.byref(st[0]), st[0].dz)
    f, _ = ctx.download_field()
    assert np.max(np.abs(f - sim.field)) < 1e-13 * np.max(np.abs(sim.field))


def test_field_track_in_drift_has_no_source(st):
    _log("field parity in a drift step (no source, pure diffraction)")
    sim = H.make_state([256] * 4, seed_power=5e3)
    ctx = H.to_gpu(sim)
    gst = H.gsteps(st)
    O.lib().orc_field_advance(C.byref(sim.s), C.byref(st[125]), st[125].dz)
    ctx.field_track(gst[125], st[125].dz)
    f, _ = ctx.download_field()
    assert np.max(np.abs(f - sim.field)) <= 1e-14 * np.max(np.abs(sim.field))


def test_space_charge_parity(st):
    _log("space charge parity: long range + short range (nz=1, nphi=1, 100 radial)")
    sim = H.make_state([1024] * 16, sc=(1, 1, 100, 0.0), longrange=True,
                       currents=cases.gaussian_current(16, 3000.0))
    ctx = H.to_gpu(sim)
    gst = H.gsteps(st)
    for i in range(3):
        O.lib().orc_beam_track(C.byref(sim.s), C.byref(st[i]), cases.GAMMA0, st[i].dz)
        ctx.beam_track(gst[i], cases.GAMMA0, st[i].dz)
    w = _beam_close(ctx, sim, 1e-12)
    _log(f"  worst rel diff {w}")


def test_slippage_and_sort_index_bit_exact(st):
    _log("slippage ring offset and sort index, bit-exact")
    sim = H.make_state([64] * 7)
    ctx = H.to_gpu(sim)
    for s in (1.0000001, 0.3, 0.6, 9.0, 0.0, 2.5, -1.2):
        O.lib().orc_slippage(C.byref(sim.s), s)
        ctx.slippage(s)
        f, first = ctx.download_field()
        assert first == sim.s.first
        assert np.array_equal(f, sim.field)
    slen = 2 * np.pi
    ctx2 = H.to_gpu(H.make_state([300, 500]))
    b = ctx2.download_beam()
    th = b["theta"]
    ref = np.zeros(len(th), np.int32)
    moved = O.lib().orc_sort_index(th.ctypes.data_as(C.POINTER(C.c_double)), len(th), slen,
                                   ref.ctypes.data_as(C.POINTER(C.c_int)))
    atar, m = ctx2.sort_index(slen)
    assert np.array_equal(atar, ref) and m == moved


def test_diag_parity():
    _log("diagnostics parity on a loaded state")
    sim = H.make_state([1024] * 8, seed_power=5e3)
    sim.s.first = 3
    ctx = H.to_gpu(sim)
    d = ctx.diag()
    r = sim.diag()
    for k in ("power", "bunching", "energy", "xsize", "ysize", "f_xsize", "nf_intensity",
              "ff_intensity"):
        e = H.rel_err(d[k], r[k], 1e-12)
        assert e < 1e-12, (k, e)
    # energyspread = sqrt(|<g^2> - <g>^2|) (Diagnostic.cpp:478) cancels
    # kappa = <g>^2 / spread^2 ~ 1e8 of the sums' relative rounding; the
    # reference's sequential sum of n terms carries ~sqrt(n) eps of it, so
    # a different (tree) order moves the result by up to sqrt(n) kappa eps.
    kappa = np.max(r["energy"] ** 2 / r["energyspread"] ** 2)
    e = H.rel_err(d["energyspread"], r["energyspread"])
    assert e < np.sqrt(1024) * kappa * 2.2e-16, (e, kappa)
    assert abs(d["global_energy"] / r["global_energy"] - 1) < 1e-13


def test_diag_aux_and_global():
    """Mean momenta, coordinate extremes, near/far-field phases and the
    global (current- and power-weighted) statistics against numpy on the
    downloaded state after 3 steps; the long- and short-range space-charge
    fields of the auxiliary block against the oracle's after the same steps."""
    _log("auxiliary and global diagnostics, 4 uneven slices, 3 steps")
    st = H.steps()
    sim = H.make_state([1024, 512, 768, 256], seed_power=5e3, longrange=True, sc=(1, 0, 100, 0.0),
                       currents=[3000.0, 2000.0, 2500.0, 1000.0])
    ctx = H.to_gpu(sim)
    ctx.run(H.gsteps(st[:3]), cases.GAMMA0, record=())
    for i in range(3):
        O.lib().orc_step_all(C.byref(sim.s), C.byref(st[i]), cases.GAMMA0)
    d = ctx.diag()
    b = ctx.download_beam()
    f, first = ctx.download_field()
    ref = H.aux_reference(b, f, first, sim.off, sim.current, sim.s.ngrid, sim.s.dgrid, cases.LAMBDA_R)
    H.check_aux(d, ref, b, 1e-10)
    # the energy spread's cancellation (see test_diag_parity)
    assert abs(d["g_energyspread"] / ref["g_energyspread"] - 1) < 1e-6
    assert H.rel_err(d["lscfield"], sim.longESC, 1e-12) < 1e-10
    assert H.rel_err(d["sscfield"], sim.sc_efield, 1e-12) < 1e-10
    assert np.array_equal(d["efield"], d["lscfield"]) and not np.any(d["wakefield"])
    ctx.close()


def test_steady_state_config0_end_to_end():
    """Config 0: steady state, one slice of 8192 (nbins 4), 151^2, 5 kW
    seed, four 125-period modules with FODO drifts (509 steps)."""
    _log("config 0 end to end (509 steps) GPU vs oracle")
    st = H.steps()
    sim = H.make_state([8192], seed_power=5e3, timerun=False)
    ctx = H.to_gpu(sim)
    g = ctx.run(H.gsteps(st), cases.GAMMA0, record=("power", "bunching"))
    r = sim.run(st, cases.GAMMA0, record=("power", "bunching"))
    ep = H.rel_err(g["power"], r["power"])
    # |b| of a quiet nbins=4 beam starts at ~1e-17: b = <exp(i theta)> is a
    # cancellation with condition number 1/|b|, so an ulp-level difference
    # in theta (device sin/cos) moves it by ~eps |theta| / sqrt(N) in
    # absolute terms.  Tolerance: 1e-10 relative plus 1e-14 absolute.
    db = np.abs(g["bunching"] - r["bunching"])
    eb = float(np.max(db / (np.abs(r["bunching"]) + 1e-4)))
    _log(f"  P(z) rel err {ep:.3e}  |b| abs err {db.max():.3e} at |b|="
         f"{r['bunching'].ravel()[np.argmax(db)]:.3e}, max |b| {r['bunching'].max():.3e}, "
         f"P_end {g['power'][-1, 0]:.6e} vs {r['power'][-1, 0]:.6e}")
    assert ep < 1e-10
    assert np.all(db <= 1e-10 * np.abs(r["bunching"]) + 1e-14), eb
    f, _ = ctx.download_field()
    assert np.max(np.abs(f - sim.field)) < 1e-10 * np.max(np.abs(sim.field))


def test_sase_time_dependent_small_end_to_end():
    """Config 1 physics on a reduced window: 24 slices x 1024, shot noise
    seed 7, long-range space charge, slippage, 509 steps on 63^2."""
    _log("time-dependent SASE, 24 slices, 509 steps, GPU vs oracle")
    st = H.steps()
    sim = H.make_state([1024] * 24, ngrid=63, gridmax=0.3e-3, longrange=True)
    ctx = H.to_gpu(sim)
    g = ctx.run(H.gsteps(st), cases.GAMMA0, record=("power", "bunching", "energy"))
    r = sim.run(st, cases.GAMMA0, record=("power", "bunching", "energy"))
    ep = H.rel_err(g["power"][1:], r["power"][1:])
    eb = H.rel_err(g["bunching"], r["bunching"])
    _log(f"  P(z) rel err {ep:.3e}  |b(z)| rel err {eb:.3e}  sum P_end {g['power'][-1].sum():.4e}")
    assert ep < 1e-10 and eb < 1e-10
    f, first = ctx.download_field()
    assert first == sim.s.first
    assert np.max(np.abs(f - sim.field)) < 1e-10 * np.max(np.abs(sim.field))


def test_config1_full_size_first_steps():
    """Config 1 at full size (1024 slices x 8192, 151^2, shot noise, LSC):
    the first three steps against the oracle (threaded CPU run)."""
    _log("config 1 full size, 3 steps GPU vs oracle (oracle multithreaded)")
    st = H.steps()
    sim = H.make_state([8192] * 1024, longrange=True, threads=16)
    ctx = H.to_gpu(sim)
    g = ctx.run(H.gsteps(st[:3]), cases.GAMMA0, record=("power", "bunching"))
    r = sim.run(st, cases.GAMMA0, record=("power", "bunching"), nsteps=3)
    ep = H.rel_err(g["power"][1:], r["power"][1:])
    eb = H.rel_err(g["bunching"], r["bunching"])
    _log(f"  P rel err {ep:.3e}  |b| rel err {eb:.3e}")
    assert ep < 1e-10 and eb < 1e-10


def test_config3_full_size_first_steps():
    """Config 3 at full size (4096 slices with the Gaussian current, 2220 to
    16384 macroparticles per slice, 40.1 M in total, 255^2, shot noise, LSC),
    the largest single-GPU workload: the first two steps against the oracle
    (threaded CPU run), P and |b| of every slice within 1e-10."""
    import g4mi_workloads as W
    _log("config 3 full size (40.1 M particles, 4096 x 255^2), 2 steps GPU vs oracle")
    cur, cnt = W.config3_profile()
    st = H.steps()
    sim = H.make_state(cnt, ngrid=255, currents=cur, longrange=True, sc=(1, 0, 100, 0.0), threads=16)
    ctx = H.to_gpu(sim)
    g = ctx.run(H.gsteps(st[:2]), cases.GAMMA0, record=("power", "bunching"))
    ctx.close()
    r = sim.run(st, cases.GAMMA0, record=("power", "bunching"), nsteps=2)
    ep = H.rel_err(g["power"][1:], r["power"][1:])
    eb = H.rel_err(g["bunching"], r["bunching"])
    _log(f"  {int(sim.off[-1])} particles: P rel err {ep:.3e}  |b| rel err {eb:.3e}")
    assert ep < 1e-10 and eb < 1e-10


def test_incoherent_and_wake_kicks(st):
    """Incoherent SR (loss and spread, beamlets of 8 over slices of 512,
    508, 512 particles: one ragged beamlet) and a per-slice wake loss after
    the push, five undulator steps and a drift: the device adds the host
    stream's kicks and eloss dz / eev exactly as the oracle (Beam.cpp:105-106);
    the wakefield diagnostic is 0 before the first step and eloss after."""
    _log("incoherent SR + wake kicks: 3 slices, 5 undulator steps + drift")
    sim = H.make_state([512, 508, 512], seed_power=5e7)
    ctx = H.to_gpu(sim)
    h = O.lib().orc_incoherent_new(4321, 0, 1, 1)
    sim.s.incoh = h
    sim.s.nbins = 8
    w = np.array([-2.0e6, 5.0e5, 1.0e6])
    sim.s.wake = w.ctypes.data_as(C.POINTER(C.c_double))
    ctx.set_incoherent(seed=4321, nbins=8, do_loss=True, do_spread=True)
    ctx.set_wake(w)
    assert np.all(ctx.diag()["wakefield"] == 0)
    gst = H.gsteps(st)
    for idx in (0, 1, 2, 3, 4, 125):
        O.lib().orc_beam_track(C.byref(sim.s), C.byref(st[idx]), cases.GAMMA0, st[idx].dz)
        ctx.beam_track(gst[idx], cases.GAMMA0, st[idx].dz)
    O.lib().orc_incoherent_free(h)
    sim.s.incoh = None
    wk = _beam_close(ctx, sim, 1e-13)
    _log(f"  worst rel diff {wk}")
    assert np.array_equal(ctx.diag()["wakefield"], w)
    ctx.close()


def test_incoherent_sase_end_to_end():
    """Whole steps with incoherent SR and wakes: 8 slices x 1024 SASE on
    63^2 over 200 steps, P(z) and |b(z)| within 1e-10 of the oracle."""
    _log("SASE with incoherent SR and wakes, 8 slices, 200 steps")
    st = H.steps()
    sim = H.make_state([1024] * 8, ngrid=63, gridmax=0.3e-3, threads=16)
    ctx = H.to_gpu(sim)
    h = O.lib().orc_incoherent_new(99, 0, 1, 1)
    sim.s.incoh = h
    sim.s.nbins = 4
    w = np.linspace(-3e5, 3e5, 8)
    sim.s.wake = w.ctypes.data_as(C.POINTER(C.c_double))
    ctx.set_incoherent(seed=99, nbins=4, do_loss=True, do_spread=True)
    ctx.set_wake(w)
    g = ctx.run(H.gsteps(st)[:200], cases.GAMMA0, record=("power", "bunching", "energy"))
    r = sim.run(st, cases.GAMMA0, record=("power", "bunching", "energy"), nsteps=200)
    O.lib().orc_incoherent_free(h)
    sim.s.incoh = None
    ep = H.rel_err(g["power"][1:], r["power"][1:])
    eb = H.rel_err(g["bunching"], r["bunching"])
    ee = H.rel_err(g["energy"], r["energy"])
    _log(f"  P(z) rel err {ep:.3e}  |b(z)| {eb:.3e}  <gamma> {ee:.3e}  "
         f"<gamma> change {g['energy'][-1].mean() - g['energy'][0].mean():.4f}")
    assert ep < 1e-10 and eb < 1e-10 and ee < 1e-12
