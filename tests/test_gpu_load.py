"""Beam loading on the device (g4mi_load_beam_device) against the oracle's
quiet loader and shot noise (LoadBeam::init, src/Loading/LoadBeam.cpp:127-212).

The device draws the same Hammersley points (integer radical inverses, the
same IEEE divisions) and evaluates the inverse erfc with the same
operations; its log/exp/sin differ from the host's by an ulp and its slice
moments are block sums, so the beams agree to rounding, not bit for bit.
The bound is 1e-12 of each coordinate's scale per slice."""
import numpy as np
import pytest

import cases
import g4mi
import oracle_py as O

pytestmark = pytest.mark.gpu

KEYS = ("gamma", "theta", "x", "y", "px", "py")


def _slices(ns):
    """Slices that exercise every branch of the loader: Twiss tilt and
    centroids, weak bunching (phase kick), strong bunching (> 0.1, the
    bisection), energy modulation, a zero-current slice."""
    out = []
    for j in range(ns):
        s = cases.slice_beam(current=cases.CURRENT * (1 + 0.2 * j))
        s.update(alphax=0.3 * (j % 3) - 0.3, alphay=-0.2 * (j % 2), xcen=1e-5 * j, ycen=-2e-6 * j,
                 pxcen=1e-3 * (j % 2), pycen=0.0)
        if j % 4 == 1:
            s.update(bunch=0.05, bunchphase=0.4)
        if j % 4 == 2:
            s.update(bunch=0.4, bunchphase=-1.0, emod=2e-3 * cases.GAMMA0, emodphase=0.7)
        if j == ns - 1:
            s.update(current=0.0)
        out.append(s)
    return out


def _oracle(counts, slices, nbins, shot, seed, islice0):
    sim = O.Sim(counts, 11, cases.GRIDMAX, cases.LAMBDA_R)
    O.load_beam(sim, slices, None, nbins, seed if shot else None, islice0=islice0, lambda0=cases.LAMBDA_R)
    return sim


def _check(got, sim, off, tol=1e-12):
    worst = 0.0
    for k in KEYS:
        ref = getattr(sim, k)
        for j in range(len(off) - 1):
            a, b = int(off[j]), int(off[j + 1])
            if b == a:
                continue
            scale = max(np.max(np.abs(ref[a:b])), 1e-300)
            e = np.max(np.abs(got[k][a:b] - ref[a:b])) / scale
            worst = max(worst, e)
            assert e < tol, (k, j, e)
    return worst


@pytest.mark.parametrize("counts,nbins,shot,islice0", [
    ([1024, 4096, 512, 2048, 0, 1024], 4, True, 0),
    ([800, 1600, 8000, 2400], 8, True, 17),
    ([3000, 999], 3, False, 5),
    ([65536, 4096], 2, True, 0),
])
def test_device_loader_matches_oracle(counts, nbins, shot, islice0):
    print(f"device loader FP64: counts {counts} nbins {nbins} shot {shot}", flush=True)
    counts = [c - c % nbins for c in counts]
    off = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)
    slices = _slices(len(counts))
    sim = _oracle(counts, slices, nbins, shot, 7, islice0)
    ctx = g4mi.Context(off, 11, cases.GRIDMAX, cases.LAMBDA_R)
    ctx.load_beam(slices, nbins, shot, 7, cases.LAMBDA_R, islice0=islice0)
    got = ctx.download_beam()
    e = _check(got, sim, off)
    print(f"  max rel err {e:.3e}", flush=True)
    ctx.close()


def test_device_loader_runs_like_an_uploaded_beam():
    """The device-loaded beam steps like the host-loaded one: 12 undulator
    steps from either agree to 1e-10 in power and bunching."""
    print("device loader: 12 steps vs an uploaded host beam", flush=True)
    counts = [2048] * 4
    off = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)
    slices = [cases.slice_beam() for _ in counts]
    host = g4mi.load_beam(off, slices, 4, True, 3, cases.LAMBDA_R)
    st = g4mi.steps_from(O.unroll(cases.fodo_elements(1, 12), cases.DELZ, cases.LAMBDA_R, cases.GAMMA0))
    field = np.zeros((len(counts), 11 * 11 * 2))
    res = []
    for dev in (False, True):
        ctx = g4mi.Context(off, 11, cases.GRIDMAX, cases.LAMBDA_R)
        if dev:
            ctx.load_beam(slices, 4, True, 3, cases.LAMBDA_R)
        else:
            ctx.upload_beam(host, [s["current"] for s in slices])
        ctx.upload_field(field.ravel(), 0)
        res.append(ctx.run(st, cases.GAMMA0, record=("power", "bunching")))
        ctx.close()
    for k in ("power", "bunching"):
        a, b = res[1][k], res[0][k]
        e = np.max(np.abs(a - b)) / np.max(np.abs(b))
        print(f"  {k} rel err {e:.3e}", flush=True)
        assert e < 1e-10, k


def test_device_loader_mixed_form_matches_upload():
    """Mixed-form contexts: the device loader's FP32 offsets equal those of
    g4mi_beam_upload of the oracle's beam to one FP32 rounding."""
    print("device loader, mixed form", flush=True)
    counts = [4096, 2048, 8192]
    off = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)
    slices = _slices(3)
    sim = _oracle(counts, slices, 4, True, 11, 2)
    out = []
    for dev in (False, True):
        ctx = g4mi.Context(off, 11, cases.GRIDMAX, cases.LAMBDA_R, precision=g4mi.FP32)
        if dev:
            ctx.load_beam(slices, 4, True, 11, cases.LAMBDA_R, islice0=2)
        else:
            ctx.upload_beam({k: getattr(sim, k) for k in KEYS}, sim.current)
        out.append(ctx.download_beam())
        ctx.close()
    for k in KEYS:
        a, b = out[1][k], out[0][k]
        # one FP32 rounding of the offset (|g - gref| ~ delgam, |th - base| < 8 pi)
        scale = np.max(np.abs(b - (np.mean(b) if k == "gamma" else 0.0)))
        e = np.max(np.abs(a - b)) / scale
        print(f"  {k} rel err {e:.3e}", flush=True)
        assert e < 2 ** -22, k


def test_device_loader_rejects_bad_nbins():
    off = np.array([0, 1022], np.int64)
    ctx = g4mi.Context(off, 11, cases.GRIDMAX, cases.LAMBDA_R)
    with pytest.raises(g4mi.G4miError, match="NBINS"):
        ctx.load_beam([cases.slice_beam()], 4, False, 1, cases.LAMBDA_R)
    ctx.close()
