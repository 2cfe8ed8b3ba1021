"""Bunching at harmonics of the beam phase (DiagBeam::getValues,
src/Core/Diagnostic.cpp:431-436, 486-500): b_h = (1/N) sum exp(i h theta)
per slice, written as Beam/bunching<h> and Beam/bunchingphase<h> for
h = 2..bunchharm (only h = bunchharm with exclusive_harmonics).  Pinned by
the textbook sum in numpy on the same particles."""
import numpy as np
import pytest

import cases
import g4mi
import harness as H


def _numpy_bunching(off, theta, hmax):
    ns = len(off) - 1
    b = np.zeros((hmax - 1, ns), complex)
    for h in range(2, hmax + 1):
        for i in range(ns):
            t = theta[off[i]:off[i + 1]]
            b[h - 2, i] = np.mean(np.exp(1j * h * t)) if len(t) else 0.0
    return b


@pytest.mark.gpu
def test_harmonic_bunching_against_numpy():
    """Harmonics 2..11 (two launches of eight and two) on a bunched SASE
    beam of three uneven slices: |b_h| within 1e-13, phases within 1e-9
    where |b_h| > 1e-6."""
    print("harmonic bunching: 3 slices, h = 2..11", flush=True)
    sim = H.make_state([512, 768, 1024], seed_power=1e6)
    ctx = H.to_gpu(sim)
    st = H.gsteps(H.steps(nund=1, nwig=60))
    ctx.run(st[:60], cases.GAMMA0)
    b, ph = ctx.diag_bunching(11)
    beam = ctx.download_beam()
    off = ctx.slices()[0]
    ref = _numpy_bunching(off, beam["theta"], 11)
    err = float(np.max(np.abs(b - np.abs(ref))))
    mask = np.abs(ref) > 1e-6
    perr = float(np.max(np.abs(np.angle(np.exp(1j * (ph[mask] - np.angle(ref[mask])))))))
    print(f"  max |b_h| {np.max(np.abs(ref)):.3e}, |b| err {err:.2e}, phase err {perr:.2e}", flush=True)
    assert err < 1e-13 and perr < 1e-9


def test_bunchharm_needs_fp64(tmp_path):
    import test_deck as T
    path = T._write(tmp_path, deck=T.EX1_IN.replace("&track\n&end", "&track\nbunchharm = 3\nprecision = fp32\n&end"),
                    root=str(tmp_path / "b"))
    with pytest.raises(RuntimeError, match="FP64"):
        g4mi.check_deck(path)


@pytest.mark.gpu
@pytest.mark.parametrize("excl", [False, True])
def test_bunchharm_deck_outputs(tmp_path, excl):
    """Example1 (63^2) with bunchharm = 3: Beam/bunching2/3 and their phases
    (only bunching3 with exclusive_harmonics), record 0 equal to numpy's sum
    over the initial beam's dump within 1e-13."""
    import test_deck as T
    print(f"deck runner: bunchharm = 3, exclusive_harmonics = {excl}", flush=True)
    root = str(tmp_path / "b")
    deck = T.EX1_IN.replace("ngrid=255", "ngrid=63").replace(
        "&track\n&end", "&write\nbeam = @.init\n&end\n\n&track\nzstop=1\nbunchharm = 3\n"
        f"exclusive_harmonics = {'true' if excl else 'false'}\n&end")
    path = T._write(tmp_path, deck=deck, root=root)
    g4mi.run_deck(path, device=0, quiet=True)
    out = root + ".out.h5"
    th = g4mi.h5_read(root + ".init.par.h5", "slice000001/theta")
    ref = _numpy_bunching(np.array([0, len(th)]), th, 3)
    b3 = g4mi.h5_read(out, "Beam/bunching3")
    assert b3.ndim == 2 and b3.shape[1] == 1 and b3.shape[0] > 2
    assert abs(b3[0, 0] - abs(ref[1, 0])) < 1e-13
    ph3 = g4mi.h5_read(out, "Beam/bunchingphase3")
    assert ph3.shape == b3.shape
    if excl:
        with pytest.raises(Exception):
            g4mi.h5_read(out, "Beam/bunching2")
    else:
        b2 = g4mi.h5_read(out, "Beam/bunching2")
        assert abs(b2[0, 0] - abs(ref[0, 0])) < 1e-13
