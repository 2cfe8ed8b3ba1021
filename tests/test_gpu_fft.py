"""GPU parity of the spectral field solver (FieldSolverFFT,
src/Core/FieldSolverFFT.cpp) against the oracle's restatement (a direct
2-D DFT per slice, pinned against numpy.fft and the analytic Gaussian
beam).  The device runs the step as E' = M E M^T + 2 S with M the circulant
1-D propagator, either as FP64 MFMA products or as FFT convolutions of
every row and column; the oracle transforms, multiplies by
exp(K2 delz) and transforms back.  The two differ only by rounding: field
within 1e-12 of max |E| per step, P(z) and |b(z)| within 1e-10 over a run."""
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


def _fft_state(counts, ngrid, gridmax=cases.GRIDMAX, filt=None, **kw):
    sim = H.make_state(counts, ngrid=ngrid, gridmax=gridmax, **kw)
    sim.s.fft_solver = 1
    if filt:
        sim.s.fft_filter = 1
        sim.s.fft_xcut, sim.s.fft_ycut, sim.s.fft_sig = filt
    ctx = H.to_gpu(sim)
    if filt:
        ctx.set_field_solver(fft=True, source_filter=True, xcut=filt[0], ycut=filt[1], sigmoid=filt[2])
    else:
        ctx.set_field_solver(fft=True)
    return sim, ctx


@pytest.mark.parametrize("variant", [0, 1, 2, 3])
@pytest.mark.parametrize("ngrid", [31, 40, 63, 151])
def test_fft_field_track_parity(ngrid, variant):
    """Deposit + spectral step on grids that take each output tile edge
    (32, 48, 64, 80) of the complex GEMM, for each GEMM variant (four or
    three real products per complex one, k chunks of 16 or 32), three
    undulator steps."""
    _log(f"spectral field step parity: 6 slices x 1024, {ngrid}^2, GEMM variant {variant}, 3 steps")
    st = H.steps()
    sim, ctx = _fft_state([1024] * 6, ngrid, seed_power=5e7)
    ctx.set_schedule("fft_variant", variant)
    ctx.set_schedule("fft_propagation", 0)
    gst = H.gsteps(st)
    for i in range(3):
        O.lib().orc_field_advance(C.byref(sim.s), C.byref(st[i]), st[i].dz)
        ctx.field_track(gst[i], st[i].dz)
    f, first = ctx.download_field()
    assert first == sim.s.first
    err = np.max(np.abs(f - sim.field)) / np.max(np.abs(sim.field))
    _log(f"  field max diff / max |E| = {err:.3e}")
    assert err < 1e-12
    assert ctx.schedule()["gemm_variant"] == variant
    assert ctx.schedule()["propagation"] == "gemm"
    ctx.close()


@pytest.mark.parametrize("ngrid", [31, 40, 64, 65, 151, 255])
def test_fft_conv_field_track_parity(ngrid):
    """The propagation as FFT convolutions (zero-padded Stockham FFTs of
    length 128, 256 or 512: radix 2, 4 and 8 passes), three undulator
    steps against the oracle's 2-D DFT step; 64 and 65 straddle the switch
    from length 128 to 256."""
    _log(f"spectral field step parity, FFT convolution: 6 slices x 1024, {ngrid}^2, 3 steps")
    st = H.steps()
    sim, ctx = _fft_state([1024] * 6, ngrid, seed_power=5e7)
    ctx.set_schedule("fft_propagation", 1)
    gst = H.gsteps(st)
    for i in range(3):
        O.lib().orc_field_advance(C.byref(sim.s), C.byref(st[i]), st[i].dz)
        ctx.field_track(gst[i], st[i].dz)
    f, first = ctx.download_field()
    assert first == sim.s.first
    err = np.max(np.abs(f - sim.field)) / np.max(np.abs(sim.field))
    sch = ctx.schedule()
    _log(f"  field max diff / max |E| = {err:.3e}, length {sch['conv_length']}")
    assert err < 1e-12
    assert sch["propagation"] == "fft-conv"
    assert sch["conv_length"] == max(128, 1 << (2 * ngrid - 2).bit_length())
    ctx.close()


def test_fft_conv_rejects_large_grids():
    sim = H.make_state([256] * 2, ngrid=301, gridmax=0.3e-3)
    ctx = H.to_gpu(sim)
    ctx.set_field_solver(fft=True)
    with pytest.raises(g4mi.G4miError, match="256"):
        ctx.set_schedule("fft_propagation", 1)
    ctx.close()


def test_fft_variant_tuner():
    """With no variant fixed, the first spectral step times all four and
    keeps the fastest; the result agrees with the oracle."""
    st = H.steps()
    sim, ctx = _fft_state([1024] * 4, 63, gridmax=0.3e-3, seed_power=5e7)
    O.lib().orc_field_advance(C.byref(sim.s), C.byref(st[0]), st[0].dz)
    ctx.field_track(H.gsteps(st)[0], st[0].dz)
    sch = ctx.schedule()
    _log(f"  tuned: {sch}")
    ms = sch["gemm_tuned_ms"]
    assert sch["gemm_variant"] in range(4) and ms[sch["gemm_variant"]] == min(ms)
    conv_faster = 0 <= sch["conv_tuned_ms"] < min(ms)
    assert sch["propagation"] == ("fft-conv" if conv_faster else "gemm")
    f, _ = ctx.download_field()
    assert np.max(np.abs(f - sim.field)) < 1e-12 * np.max(np.abs(sim.field))


def test_fft_drift_step_is_pure_diffraction():
    """A drift step: no source, the field is only propagated."""
    st = H.steps()
    sim, ctx = _fft_state([256] * 4, 63, gridmax=0.3e-3, seed_power=5e3)
    O.lib().orc_field_advance(C.byref(sim.s), C.byref(st[125]), st[125].dz)
    ctx.field_track(H.gsteps(st)[125], st[125].dz)
    f, _ = ctx.download_field()
    err = np.max(np.abs(f - sim.field)) / np.max(np.abs(sim.field))
    _log(f"  drift step: field max diff / max |E| = {err:.3e}")
    assert err < 1e-13


@pytest.mark.parametrize("prop", [0, 1])
def test_fft_source_filter_parity(prop):
    """Source filtered by a sigmoid at a third of the band (xcut = ycut =
    0.3, width 0.05), two undulator steps, propagation by the GEMMs (0) or
    by FFT convolution (1)."""
    _log(f"spectral step with source filter: 4 slices x 1024, 63^2, propagation {prop}")
    st = H.steps()
    sim, ctx = _fft_state([1024] * 4, 63, gridmax=0.3e-3, seed_power=5e7, filt=(0.3, 0.3, 0.05))
    ctx.set_schedule("fft_propagation", prop)
    gst = H.gsteps(st)
    for i in range(2):
        O.lib().orc_field_advance(C.byref(sim.s), C.byref(st[i]), st[i].dz)
        ctx.field_track(gst[i], st[i].dz)
    f, _ = ctx.download_field()
    err = np.max(np.abs(f - sim.field)) / np.max(np.abs(sim.field))
    _log(f"  field max diff / max |E| = {err:.3e}")
    assert err < 1e-12


@pytest.mark.parametrize("prop", [0, 1])
def test_fft_sase_end_to_end(prop):
    """Time-dependent SASE with slippage on the spectral solver: 12 slices
    x 1024, 63^2, the first 200 steps (two undulator modules and a FODO
    drift), P(z) and |b(z)| within 1e-10; propagation by the GEMMs (0) or
    by FFT convolution (1)."""
    _log(f"spectral solver SASE, 12 slices, 200 steps, GPU vs oracle, propagation {prop}")
    st = H.steps()
    sim, ctx = _fft_state([1024] * 12, 63, gridmax=0.3e-3, longrange=True, threads=16)
    ctx.set_schedule("fft_propagation", prop)
    g = ctx.run(H.gsteps(st)[:200], cases.GAMMA0, record=("power", "bunching"))
    r = sim.run(st, cases.GAMMA0, record=("power", "bunching"), nsteps=200)
    ep = H.rel_err(g["power"][1:], r["power"][1:])
    eb = H.rel_err(g["bunching"], r["bunching"])
    _log(f"  P(z) rel err {ep:.3e}  |b(z)| rel err {eb:.3e}  sum P_end {g['power'][-1].sum():.4e}")
    assert ep < 1e-10 and eb < 1e-10
    f, first = ctx.download_field()
    assert first == sim.s.first
    assert np.max(np.abs(f - sim.field)) < 1e-10 * np.max(np.abs(sim.field))


def test_fft_solver_rejected_in_mixed_form():
    sim = H.make_state([256] * 2, ngrid=31, gridmax=0.3e-3)
    ctx = H.to_gpu(sim, precision=g4mi.FP32)
    with pytest.raises(g4mi.G4miError) as e:
        ctx.set_field_solver(fft=True)
    assert "FP64" in str(e.value)


def _tilted_fields(ng, gridmax, lam, w0, tilts):
    x = np.linspace(-gridmax, gridmax, ng)
    X, Y = np.meshgrid(x, x)  # [iy][ix]
    k = 2 * np.pi / lam
    out = []
    for tx, ty, amp in tilts:
        e = amp * np.exp(-(X ** 2 + Y ** 2) / w0 ** 2) * np.exp(1j * k * (tx * X + ty * Y))
        out.append(np.stack([e.real, e.imag], axis=-1).reshape(-1, 2))
    return np.ascontiguousarray(np.array(out))


def _numpy_spectral(field, ng, first, lam, dgrid):
    """DiagField's spectral block restated with numpy.fft: moments of
    |DFT(E)|^2 at the grid offset of each FFT index."""
    ns = field.shape[0]
    q = (np.arange(ng) - (ng + 1) // 2) % ng - 0.5 * (ng - 1)  # offset of FFT index k
    scl = lam / ng / dgrid
    res = {k: np.zeros(ns) for k in ("xdivergence", "ydivergence", "xpointing", "ypointing")}
    tot = np.zeros(5)
    for k in range(ns):
        e = (field[k, :, 0] + 1j * field[k, :, 1]).reshape(ng, ng)
        w = np.abs(np.fft.fft2(e)) ** 2
        m = np.array([w.sum(), (w * q[None, :]).sum(), (w * q[None, :] ** 2).sum(),
                      (w * q[:, None]).sum(), (w * q[:, None] ** 2).sum()])
        tot += m
        v = m[1:] / m[0]
        i = (ns + k - first) % ns
        res["xpointing"][i] = v[0] * scl
        res["xdivergence"][i] = np.sqrt(abs(v[1] - v[0] ** 2)) * scl
        res["ypointing"][i] = v[2] * scl
        res["ydivergence"][i] = np.sqrt(abs(v[3] - v[2] ** 2)) * scl
    v = tot[1:] / tot[0]
    res["g_xpointing"] = v[0] * scl
    res["g_xdivergence"] = np.sqrt(abs(v[1] - v[0] ** 2)) * scl
    res["g_ypointing"] = v[2] * scl
    res["g_ydivergence"] = np.sqrt(abs(v[3] - v[2] ** 2)) * scl
    return res


@pytest.mark.parametrize("prec,ng,dft", [(g4mi.FP64, 121, "bluestein"), (g4mi.FP32, 121, "bluestein"),
                                         (g4mi.FP64, 255, "bluestein"), (g4mi.FP64, 121, "gemm")])
def test_spectral_diagnostics(prec, ng, dft):
    """Divergence and pointing (Diagnostic.cpp:789-846, globals :952-963)
    of three tilted Gaussian beams (w0 = 30 um on 121^2 cells of 5 um or
    255^2 of 2.4 um, ring offset first = 1), the 2-D DFT by Bluestein's
    chirp convolution (zero-padded FFTs of length 256 or 512) or by the
    dense DFT products: against the same moments from numpy.fft (1e-12 in
    the FP64 form, 1e-5 in the FP32 form) and against optics: the pointing
    is the tilt and the rms divergence lambda / (2 pi w0), to 1e-6."""
    gm, w0 = 0.3e-3, 30e-6
    lam = cases.LAMBDA_R
    sig = lam / (2 * np.pi * w0)
    tilts = [(3 * sig, -2 * sig, 1e7), (0.0, 0.0, 2e7), (-1.5 * sig, 4 * sig, 5e6)]
    _log(f"spectral diagnostics ({'FP64' if prec == g4mi.FP64 else 'FP32'} form), 3 tilted beams, {ng}^2, {dft}")
    sim = H.make_state([64] * 3, ngrid=ng, gridmax=gm)
    sim.field[:] = _tilted_fields(ng, gm, lam, w0, tilts)
    sim.s.first = 1
    ctx = H.to_gpu(sim, precision=prec)
    if dft == "gemm":
        ctx.set_schedule("fft_propagation", 0)
    d = ctx.diag(spectral=True)
    ref = _numpy_spectral(sim.field, ng, 1, lam, 2 * gm / (ng - 1))
    tol = 1e-12 if prec == g4mi.FP64 else 1e-5
    for k in ("xdivergence", "ydivergence", "xpointing", "ypointing"):
        np.testing.assert_allclose(d[k], ref[k], rtol=tol, atol=tol * sig, err_msg=k)
        np.testing.assert_allclose(d["g_" + k], ref["g_" + k], rtol=tol, atol=tol * sig, err_msg=k)
    # optics: output slice i holds storage slice (i + first) mod ns
    for i in range(3):
        tx, ty, _ = tilts[(i + 1) % 3]
        assert abs(d["xpointing"][i] - tx) < 1e-6 * sig and abs(d["ypointing"][i] - ty) < 1e-6 * sig
        assert abs(d["xdivergence"][i] / sig - 1) < 1e-6 and abs(d["ydivergence"][i] / sig - 1) < 1e-6
    _log(f"  divergence {d['xdivergence']} (optics {sig:.6e}); pointing {d['xpointing']}")
    ctx.close()
